#!/usr/bin/env python
"""Benchmark of the MI355X finite element assembly path.

Workload (one "step"): Mooney-Rivlin residual + FP64 analytic Jacobian
assembly of the 32x32x32 triquadratic/trilinear unit-cube mesh (32,768
elements, 859,812 DOFs, 178,723,696 Jacobian nonzeros), interior nodes
jittered by U[-0.1h, 0.1h], c10 = 2, c01 = 6, at a deformed state (10%
uniaxial stretch plus a seeded perturbation).  Inputs are resident in HBM
before the timed region.  Prints one JSON line.

Also reported (``secondary``): the FP32-Jacobian variant; time to a 1e-10
relative residual per load step on the 32^3 case (20% stretch in 10 load
steps) for plain and learned starts with FP64 and FP32 Jacobians, and the
same on the 16^3 case (10% in 10 steps); the solver-matrix assembly; Laplace
assembly throughput at 8^3 / 64^3 / 128^3 trilinear; and the batched
workload (256 meshes of uneven size assembled in one launch per kernel).
``cpu_baseline`` times the C oracle (an FP64 restatement of the reference
element loop with the reference's per-entry CSR bisection) on this host,
single thread, on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_MFMA_PEAK_TFLOPS = 157.3
# Algorithmic FLOP per element of the re-associated Jacobian contraction the
# element-matrix kernels compute (K_ab = sum_g D_g^T W_ab,g D_g over the six
# lower 27x27 blocks, diagonal blocks as triangles, plus the pressure rows):
#   D_g (27*81*3 MAC) 13,122 + W (27*(54 mul + 27 add)) 2,187
#   + U = D^T W (6 blocks * 27*27*3*3 MAC) 78,732
#   + U D (3 * 27*27*81 + 3 * 378*81 MAC) 538,002
#   + pressure rows (8*81*27 MAC + 216 mul) 35,208            = 667,251
# (the reference element loop does about 0.68 MFLOP per element).
MR_JACOBIAN_FLOP_PER_ELEMENT = 667251
# MFMA work issued per element: FP64 1659 mfma_f64_4x4x4 (512 FLOP each),
# FP32 483 mfma_f32_16x16x4 (2048 FLOP each), padding included
MR_JACOBIAN_ISSUED_FLOP_F64 = 1659 * 512
MR_JACOBIAN_ISSUED_FLOP_F32 = 483 * 2048


def deformed_state(eq, mesh, stretch=1.10, seed=17):
    """Homogeneous uniaxial stretch plus a small seeded perturbation."""
    nq, nl = mesh.n_nodes, mesh.n_nodes_l
    X = mesh.coords
    lam = np.array([stretch, stretch ** -0.5, stretch ** -0.5])
    rng = np.random.default_rng(seed)
    z = X * lam[:, None] + 0.002 * rng.uniform(-1, 1, X.shape)
    p = 0.5 * rng.standard_normal(nl)
    return np.concatenate([z.ravel(), p])


def laplace_bytes(n_elements, n_nodes, nnz):
    # connectivity int32x8 + coordinates 3 doubles + CSR values + RHS
    return 32 * n_elements + 24 * n_nodes + 8 * nnz + 8 * n_nodes


def traffic_from_profile(kernel, workload):
    """(HBM bytes per launch, source file): measured with rocprofv3 --pmc
    (FETCH_SIZE and WRITE_SIZE, scaled by a copy-kernel calibration) in a
    separate profiling run and kept under profiles/; (None, None) if that
    kernel and workload have not been measured."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload}.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch"), os.path.relpath(path, ROOT)


def time_steps(fn, steps, warmup, cm, only=None):
    """Times `steps` calls of fn between two events; kernel timers on (all
    kernels, or only kernel `only`: one event pair per launch of it)."""
    for _ in range(warmup):
        fn()
    cm.synchronise()
    cm.kernel_timing_enable(True, only=only)
    cm.event_record(0)
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    cm.event_record(1)
    cm.synchronise()
    wall = time.perf_counter() - t0
    ms = cm.event_elapsed(0, 1)
    return ms, wall


def laplace_line(cm, n, steps):
    """Laplace assembly of an n^3 structured grid: the fused one-pass kernel
    (the default for such grids) and the element buffer + row gather."""
    mesh = cm.RegularMesh.generate(1, n)
    eq = cm.LaplaceEquationsSet(mesh)
    nbytes = laplace_bytes(mesh.n_elements, mesh.n_nodes, eq.nnz)
    out = {"n": n, "elements": mesh.n_elements, "nnz": eq.nnz, "algorithmic_bytes": nbytes}
    for form, fused, names in (("fused", 2, ("laplace_fused",)),
                               ("two_pass", 0, ("laplace_element", "laplace_row_gather"))):
        eq.set_fused(fused)
        ms, _ = time_steps(eq.assemble, steps, 2, cm)
        kms = sum(cm.kernel_timing(k)[0] for k in names)
        fms, _ = cm.kernel_timing("fill")
        cm.kernel_timing_enable(False)
        r = {"elements_per_s": mesh.n_elements * steps / (ms / 1e3), "ms_per_assembly": ms / steps,
             "kernel_ms_per_assembly": (kms + fms) / steps,
             "achieved_GBs": nbytes / ((kms + fms) / steps / 1e3) / 1e9}
        r["hbm_frac"] = r["achieved_GBs"] / HBM_PEAK_GBS
        out[form] = r
    # the default form: fused where the grid fills the GPU
    out["default_form"] = "fused" if eq.set_fused(1) else "two_pass"
    out.update(out[out["default_form"]])
    eq.destroy()
    return out


def solver_matrix_line(cm, eq, steps):
    """Solver matrix and vector of the MR set under the uniaxial conditions:
    the FP64 Jacobian restricted to the free DOFs, and the residual of the free
    rows (SOLVER_MATRICES_STATIC_ASSEMBLE).  HBM-bound compaction."""
    fixed = eq.uniaxial_boundary_conditions(1.1)
    eq.residual_evaluate()
    eq.jacobian_evaluate(cm.OC_FP64)
    eq.solver_matrices_assemble(cm.OC_FP64)  # first call builds the structure
    ms, _ = time_steps(lambda: eq.solver_matrices_assemble(cm.OC_FP64), steps, 2, cm, only="solver_matrix_fp64")
    kms, launches = cm.kernel_timing("solver_matrix_fp64")
    cm.kernel_timing_enable(False)
    n, s_nnz = eq.solver_matrix_info()
    rp, _ = eq.matrix_structure(0)
    free = np.ones(eq.n_rows, bool)
    free[fixed] = False
    lens = np.diff(rp.astype(np.int64))
    nnz_free_rows = int(lens[free].sum())
    # column indices of the free rows, the free values read and written, the
    # DOF -> solver map once, row metadata (two offsets, the row, the solver
    # offset) and the vector (residual read, solver vector written)
    nbytes = 4 * nnz_free_rows + 16 * s_nnz + 4 * eq.n_rows + 28 * n + 16 * n
    avg_s = kms / launches / 1e3
    return {
        "workload": "16^3 Mooney-Rivlin FP64 Jacobian, uniaxial Dirichlet conditions",
        "solver_rows": n, "solver_nnz": s_nnz, "equations_nnz": eq.nnz,
        "ms_per_assembly": ms / steps, "kernel_ms_per_assembly": avg_s * 1e3,
        "algorithmic_bytes": nbytes, "achieved_GBs": nbytes / avg_s / 1e9,
        "hbm_frac": nbytes / avg_s / 1e9 / HBM_PEAK_GBS,
    }


def newton_line(cm, n, stretch, load_steps, krylov_rtol):
    """Time to 1e-10 x the plain start's residual per load step (uniaxial
    stretch in `load_steps` steps): plain start (previous solution + boundary
    increment) and learned warm start, each with FP64 and FP32 Jacobians, all
    four with the same forcing rule and preconditioner."""
    from cm_amd.train_warm_start import DEFAULT_WEIGHTS, solve_steps
    from cm_amd.warm_start import WarmStartMLP
    case = dict(n=n, c10=2.0, c01=6.0, stretch=stretch, seed=2024)
    net = WarmStartMLP.load(DEFAULT_WEIGHTS)
    runs = {
        "plain_fp64": solve_steps(case, load_steps, krylov_rtol=krylov_rtol),
        "plain_fp32": solve_steps(case, load_steps, precision=cm.OC_FP32, krylov_rtol=krylov_rtol),
        "learned_fp64": solve_steps(case, load_steps, net=net, krylov_rtol=krylov_rtol),
        "learned_fp32": solve_steps(case, load_steps, net=net, precision=cm.OC_FP32, krylov_rtol=krylov_rtol),
    }
    net.destroy()

    def summary(rows):
        return {"converged_steps": sum(r["converged"] for r in rows), "steps": len(rows),
                "seconds_per_step": float(np.mean([r["seconds"] for r in rows])),
                "seconds": [round(r["seconds"], 4) for r in rows],
                "newton_iterations": [r["newton"] for r in rows], "krylov_iterations": [r["krylov"] for r in rows],
                "start_residual_over_plain_start": [r["start_norm"] / r["reference_norm"] for r in rows],
                "final_residual_over_plain_start": [r["final_norm"] / r["reference_norm"] for r in rows]}
    out = {"workload": f"{n}^3 Mooney-Rivlin, c10=2, c01=6, {round(100 * stretch)}% stretch in {load_steps} load "
                       "steps, each step to 1e-10 x the plain start's FP64 residual; GMRES(100), block-triangular "
                       "multigrid preconditioner, "
                       + ("Eisenstat-Walker forcing" if krylov_rtol <= 0 else f"forcing rtol {krylov_rtol:g}")
                       + ", linear solves floored at half the Newton target"}
    out.update({k: summary(v) for k, v in runs.items()})
    base = out["plain_fp64"]["seconds_per_step"]
    out["speedup_vs_plain_fp64"] = {k: base / out[k]["seconds_per_step"] for k in runs}
    return out


def batched_line(cm, n_meshes, reps=3):
    """Batched meshes (n ~ U{4..16}, seeded materials and stretches): residual
    + Jacobian assembly of every mesh in one launch per kernel, FP32 and FP64,
    after the launch-shape autotuner; and the same Jacobian replayed as a graph."""
    from cm_amd.batch_workload import build, draw_cases
    cases = draw_cases(n_meshes)
    meshes, batch, setup = build(cases)
    batch.load_increment_apply(1, 5)
    E = batch.n_elements
    out = {"workload": f"{n_meshes} Mooney-Rivlin meshes, n ~ U{{4..16}}, c10 ~ U[1,5], c01 ~ U[0,10], jitter 0.1h, "
                       "first load step state", "meshes": n_meshes, "elements": E, "dofs": batch.n_rows,
           "nnz": batch.nnz, "setup_s": setup}
    for prec, name in ((cm.OC_FP32, "fp32"), (cm.OC_FP64, "fp64")):
        batch.jacobian_evaluate(prec)
        out[f"autotune_{name}"] = batch.autotune(prec)
        ms, _ = time_steps(lambda: (batch.residual_evaluate(), batch.jacobian_evaluate(prec)), reps, 1, cm)
        em, _ = cm.kernel_timing(f"mr_element_matrix_{name}")
        ks = {k: cm.kernel_timing(k)[0] / reps for k in
              ("mr_gauss_jacobian", f"mr_element_matrix_{name}", f"csr_row_gather_{name}", "mr_residual", "dof_gather")}
        cm.kernel_timing_enable(False)
        gms, _ = time_steps(lambda: batch.jacobian_evaluate_graph(prec), reps, 1, cm)
        cm.kernel_timing_enable(False)
        peak = FP32_MFMA_PEAK_TFLOPS if name == "fp32" else FP64_PEAK_TFLOPS
        tflops = MR_JACOBIAN_FLOP_PER_ELEMENT * E / (em / reps / 1e3) / 1e12  # one launch per step
        out[name] = {"elements_per_s": E * reps / (ms / 1e3), "ms_per_residual_and_jacobian": ms / reps,
                     "graph_jacobian_ms": gms / reps, "kernel_ms": ks,
                     "element_matrix_TFLOPs": tflops, "element_matrix_frac": tflops / peak}
    # one load step of every mesh: learned warm start, FP32 Jacobians, FP64
    # residual, multigrid, the Newton-step graph; 1e-10 x each plain start
    from cm_amd.train_warm_start import DEFAULT_WEIGHTS, solver_options
    from cm_amd.warm_start import WarmStartMLP
    net = WarmStartMLP.load(DEFAULT_WEIGHTS)
    batch.set_multigrid()
    batch.set_dofs(batch.undeformed_dofs())
    batch.load_increment_apply(1, 5)
    refs = batch.mesh_residual_norms()
    batch.set_dofs(batch.undeformed_dofs())
    cm.synchronise()
    t0 = time.perf_counter()
    net.warm_start(batch, np.array([c["stretch"] / 5 for c in cases]))
    batch.load_increment_apply(1, 5)
    st, finals = batch.newton_solve(solver_options(cm.OC_FP32), reference_norms=refs, raise_on_failure=False)
    cm.synchronise()
    out["learned_load_step"] = {"seconds": time.perf_counter() - t0, "converged_meshes": int(np.sum(finals <= 1e-10 * refs)),
                                "meshes": n_meshes, "newton_iterations": st["newton_iterations"],
                                "krylov_iterations": st["krylov_iterations"],
                                "max_final_over_reference": float(np.max(finals / refs))}
    # load step 2 of every mesh from the step-1 solution, once from the plain
    # start with FP64 Jacobians and once from the learned start with FP32
    # Jacobians, with the forcing rule of newton_per_load_step; both after the
    # set's one-time solver setup
    from cm_amd.train_warm_start import PLAIN_KRYLOV_RTOL
    d1 = batch.get_dofs()
    batch.load_increment_apply(2, 5)
    refs2 = batch.mesh_residual_norms()
    runs = {}
    for name, prec, warm in (("plain_fp64", cm.OC_FP64, False), ("learned_fp32", cm.OC_FP32, True)):
        batch.set_dofs(d1)
        cm.synchronise()
        t0 = time.perf_counter()
        if warm:
            net.warm_start(batch, np.array([c["stretch"] / 5 for c in cases]))
        batch.load_increment_apply(2, 5)
        st, finals = batch.newton_solve(solver_options(prec, krylov_rtol=PLAIN_KRYLOV_RTOL), reference_norms=refs2,
                                        raise_on_failure=False)
        cm.synchronise()
        runs[name] = {"seconds": time.perf_counter() - t0, "converged_meshes": int(np.sum(finals <= 1e-10 * refs2)),
                      "newton_iterations": st["newton_iterations"], "krylov_iterations": st["krylov_iterations"],
                      "max_final_over_reference": float(np.max(finals / refs2))}
    runs["speedup"] = runs["plain_fp64"]["seconds"] / runs["learned_fp32"]["seconds"]
    out["load_step_2"] = runs
    net.destroy()
    batch.destroy()
    return out


def cpu_baseline(mesh, dofs, c10, c01, seconds):
    """Oracle (C FP64 restatement of the reference loop) on this host: one
    thread in the reference's element order, and all cores (up to 16 threads,
    element colours over OpenMP threads) on the same bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc
    rp, ci = orc.pattern([(mesh.n_nodes, mesh.elem_nodes)] * 3 + [(mesh.n_nodes_l, mesh.elem_nodes_l)])
    reps, t_total = 0, 0.0
    while t_total < seconds or reps == 0:
        t0 = time.perf_counter()
        orc.mr_residual(mesh.elem_nodes, mesh.elem_nodes_l, mesh.coords, dofs, c10, c01)
        orc.mr_jacobian(mesh.elem_nodes, mesh.elem_nodes_l, mesh.coords, dofs, c10, c01, rp, ci)
        t_total += time.perf_counter() - t0
        reps += 1
    threads = max(1, min(os.cpu_count() or 1, 16))
    preps, p_total = 0, 0.0
    while p_total < seconds / 4 or preps == 0:
        t0 = time.perf_counter()
        orc.mr_assemble_parallel(mesh.elem_nodes, mesh.elem_nodes_l, mesh.coords, dofs, c10, c01, rp, ci, threads)
        p_total += time.perf_counter() - t0
        preps += 1
    return {
        "value": mesh.n_elements * reps / t_total,
        "unit": "elements/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{mesh.n[0]}^3 MR residual + FP64 analytic Jacobian with per-entry CSR bisection add "
                  f"({mesh.n_elements} elements) x {reps} repetitions, {t_total:.1f} s, setup excluded",
        "all_cores": {"value": mesh.n_elements * preps / p_total, "threads": threads,
                      "sample": f"same workload, element colours over OpenMP threads, x {preps} repetitions, "
                                f"{p_total:.1f} s"},
    }


def progress(msg):
    print(msg, file=sys.stderr, flush=True)


def kernel_roofline(name, kind, per_launch, avg_launch_s, traffic_key, workload_key):
    """Roofline object of one kernel from its event-timed average launch."""
    traffic, source = traffic_from_profile(traffic_key, workload_key)
    if kind == "mfma64" or kind == "mfma32":
        peak = FP64_PEAK_TFLOPS if kind == "mfma64" else FP32_MFMA_PEAK_TFLOPS
        achieved = MR_JACOBIAN_FLOP_PER_ELEMENT * per_launch / avg_launch_s / 1e12
        issued = (MR_JACOBIAN_ISSUED_FLOP_F64 if kind == "mfma64" else MR_JACOBIAN_ISSUED_FLOP_F32) * per_launch
        out = {"kernel": name, "bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
               "frac": achieved / peak, "traffic": traffic,
               "algorithmic_count": f"{MR_JACOBIAN_FLOP_PER_ELEMENT} FLOP per element x {per_launch} elements",
               "issued_mfma_TFLOPs": issued / avg_launch_s / 1e12,
               "issued_mfma_frac": issued / avg_launch_s / 1e12 / peak}
    else:  # hbm: per_launch is the algorithmic bytes of the launch
        achieved = per_launch / avg_launch_s / 1e9
        out = {"kernel": name, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "algorithmic_count": f"{per_launch} bytes"}
    out["avg_launch_ms"] = avg_launch_s * 1e3
    out["traffic_source"] = (f"{source}: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of a separate profiling run, "
                             "scaled by a copy-kernel calibration") if source else None
    return out


def gather_bytes(eq, value_bytes):
    """Algorithmic bytes of the scatter into CSR: the Jacobian's CSR values,
    written once (the element matrices the row gather reads back are this
    implementation's intermediate, not part of the count)."""
    return eq.nnz * value_bytes


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=32, help="elements per direction of the MR workload")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--no-newton", action="store_true")
    ap.add_argument("--load-steps", type=int, default=10)
    ap.add_argument("--newton-rtol", type=float, default=1e-5,
                    help="GMRES forcing of every Newton step in the load-step comparison (<= 0: Eisenstat-Walker)")
    ap.add_argument("--batched-meshes", type=int, default=256, help="0 skips the batched-meshes line")
    ap.add_argument("--profile", choices=["jacobian", "jacobian32", "laplace", "laplace2", "copy", "newton"],
                    default=None,
                    help="run only one kernel sequence for rocprofv3: the FP64 (or FP32) Jacobian assembly of the "
                         "MR workload (--steps times), the 128^3 Laplace assembly (fused; laplace2: element buffer "
                         "+ row gather), a device copy of 512 MiB (the "
                         "FETCH_SIZE/WRITE_SIZE calibration), or --steps load steps of the --n case (20%% in 10 "
                         "steps) with --profile-start learned or plain")
    ap.add_argument("--profile-start", choices=["learned", "plain"], default="learned")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU sharding of this "
              "path is a design (element partitions with halo exchange) and is not run here.", file=sys.stderr)
        sys.exit(2)

    if args.profile == "copy":
        import torch
        a = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(args.steps):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"profile": "copy", "bytes_read_per_launch": a.numel() * 8,
                          "bytes_written_per_launch": b.numel() * 8, "launches": args.steps}))
        return

    import cm_amd as cm
    cm.initialise(0)
    if args.profile in ("laplace", "laplace2"):
        mesh = cm.RegularMesh.generate(1, 128)
        eq = cm.LaplaceEquationsSet(mesh)
        eq.set_fused(1 if args.profile == "laplace" else 0)
        for _ in range(args.steps):
            eq.assemble()
        cm.synchronise()
        print(json.dumps({"profile": args.profile, "steps": args.steps, "elements": mesh.n_elements, "nnz": eq.nnz,
                          "algorithmic_bytes": laplace_bytes(mesh.n_elements, mesh.n_nodes, eq.nnz)}))
        return
    if args.profile == "newton":
        from cm_amd.train_warm_start import DEFAULT_WEIGHTS, solve_steps
        from cm_amd.warm_start import WarmStartMLP
        case = dict(n=args.n, c10=2.0, c01=6.0, stretch=0.20, seed=2024)
        net = WarmStartMLP.load(DEFAULT_WEIGHTS) if args.profile_start == "learned" else None
        rows = solve_steps(case, args.load_steps, net=net, precision=cm.OC_FP32 if net else cm.OC_FP64,
                           max_steps=args.steps, krylov_rtol=args.newton_rtol, log=progress)
        print(json.dumps({"profile": "newton", "start": args.profile_start, "n": args.n, "steps": rows}))
        return
    c10, c01 = 2.0, 6.0
    mesh = cm.RegularMesh.generate(2, args.n, jitter=0.1, seed=2024)
    eq = cm.MooneyRivlinEquationsSet(mesh, c10, c01)
    dofs = deformed_state(eq, mesh)
    eq.set_dofs(dofs)

    def step():
        eq.residual_evaluate()
        eq.jacobian_evaluate(cm.OC_FP64)

    if args.profile in ("jacobian", "jacobian32"):
        prec = cm.OC_FP64 if args.profile == "jacobian" else cm.OC_FP32
        for _ in range(args.steps):
            eq.jacobian_evaluate(prec)
        cm.synchronise()
        print(json.dumps({"profile": args.profile, "steps": args.steps, "elements": mesh.n_elements}))
        return

    E = mesh.n_elements
    wl = f"mr{args.n}"
    # untimed pass with every kernel timed: which kernel dominates a step
    progress(f"MR {args.n}^3: {E} elements, {eq.n_rows} DOFs, {eq.nnz} nonzeros")
    time_steps(step, args.steps, args.warmup, cm)
    pre = {k: cm.kernel_timing(k)[0] for k in ("mr_element_matrix_fp64", "csr_row_gather_fp64")}
    cm.kernel_timing_enable(False)
    dominant = max(pre, key=pre.get)
    # timed region: only the dominant kernel carries events (one pair per launch);
    # the per-kernel breakdown comes from a second pass with every kernel timed
    ms, wall = time_steps(step, args.steps, 0, cm, only=dominant)
    dom_ms, dom_launches = cm.kernel_timing(dominant)
    cm.kernel_timing_enable(False)
    time_steps(step, args.steps, 0, cm)
    kernels = {name: cm.kernel_timing(name) for name in
               ("mr_gauss_jacobian", "mr_element_matrix_fp64", "csr_row_gather_fp64", "mr_residual", "dof_gather")}
    cm.kernel_timing_enable(False)
    value = E * args.steps / (ms / 1e3)
    em, el = kernels["mr_element_matrix_fp64"]
    gm, gl = kernels["csr_row_gather_fp64"]
    # algorithmic count per launch: the step's count over the launches per step
    rl_elem = kernel_roofline("mr_element_matrix_f64_kernel", "mfma64", E * args.steps / el,
                              (dom_ms / dom_launches if dominant == "mr_element_matrix_fp64" else em / el) / 1e3,
                              "mr_element_matrix_fp64", wl)
    rl_gather = kernel_roofline("csr_row_gather_kernel<double>", "hbm", gather_bytes(eq, 8) * args.steps / gl,
                                (dom_ms / dom_launches if dominant == "csr_row_gather_fp64" else gm / gl) / 1e3,
                                "csr_row_gather_fp64", wl)
    roofline = rl_elem if dominant == "mr_element_matrix_fp64" else rl_gather
    other = rl_gather if dominant == "mr_element_matrix_fp64" else rl_elem
    # the residual against both of its roofs (its algorithmic counts: about
    # 3.2 kFLOP per Gauss point x 27, and per element the 27 x 6 + 8 nodal
    # values read, 35 int32 gather indices; per DOF the residual written)
    rm, rl = kernels["mr_residual"]
    res_s = rm / rl / 1e3
    res_flop, res_bytes = 3200.0 * 27 * E, (27 * 6 + 8) * 8.0 * E + 35 * 4.0 * E + 8.0 * eq.n_rows
    rl_res = {"kernel": "mr_residual_element_kernel", "avg_launch_ms": rm / rl,
              "fp64": {"achieved": res_flop / res_s / 1e12, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                       "frac": res_flop / res_s / 1e12 / FP64_PEAK_TFLOPS, "algorithmic_count": f"3200 FLOP x 27 GP x {E} elements"},
              "hbm": {"achieved": res_bytes / res_s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": res_bytes / res_s / 1e9 / HBM_PEAK_GBS, "algorithmic_count": f"{res_bytes:.0f} bytes"}}
    secondary = {
        "kernel_ms_per_step": {name: v[0] / args.steps for name, v in kernels.items()},
        "kernel_launches_per_step": {name: v[1] / args.steps for name, v in kernels.items()},
        "wall_ms_per_step": wall * 1e3 / args.steps,
        "roofline_other_assembly_kernel": other,
        "roofline_residual": rl_res,
    }
    progress(f"assembly: {value:.4g} elements/s, {ms / args.steps:.3f} ms per step; dominant {dominant}")
    if not args.no_secondary:
        eq.jacobian_evaluate(cm.OC_FP32)
        ms32, _ = time_steps(lambda: eq.jacobian_evaluate(cm.OC_FP32), args.steps, 1, cm)
        j32, l32 = cm.kernel_timing("mr_element_matrix_fp32")
        g32, k32 = cm.kernel_timing("csr_row_gather_fp32")
        cm.kernel_timing_enable(False)
        v64, v32 = eq.matrix_values(), eq.matrix_values_sp().astype(np.float64)
        secondary["mr_jacobian_fp32"] = {
            "error_vs_fp64_rel_frobenius": float(np.linalg.norm(v32 - v64) / np.linalg.norm(v64)),
            "elements_per_s": E * args.steps / (ms32 / 1e3),
            "element_matrix": kernel_roofline("mr_element_matrix_kernel<float>", "mfma32", E * args.steps / l32,
                                              j32 / l32 / 1e3, "mr_element_matrix_fp32", wl),
            "row_gather": kernel_roofline("csr_row_gather_kernel<float>", "hbm", gather_bytes(eq, 4) * args.steps / k32,
                                          g32 / k32 / 1e3, "csr_row_gather_fp32", wl),
            "kernel_ms_per_step": {"mr_element_matrix_fp32": j32 / args.steps, "csr_row_gather_fp32": g32 / args.steps},
        }
        del v64, v32
        progress("FP32 Jacobian done")
    if not args.no_newton:
        progress(f"load steps, {args.n}^3, 20% in {args.load_steps} steps")
        secondary["newton_per_load_step"] = newton_line(cm, args.n, 0.20, args.load_steps, args.newton_rtol)
        if not args.no_secondary:
            progress("load steps, 16^3, 10% in 10 steps")
            secondary["newton_per_load_step_16"] = newton_line(cm, 16, 0.10, 10, args.newton_rtol)
    if not args.no_secondary:
        m16 = cm.RegularMesh.generate(2, 16, jitter=0.1, seed=2024)
        e16 = cm.MooneyRivlinEquationsSet(m16, c10, c01)
        e16.set_dofs(deformed_state(e16, m16))
        secondary["solver_matrix"] = solver_matrix_line(cm, e16, 20)
        e16.destroy()
        secondary["laplace"] = [laplace_line(cm, n, s) for n, s in ((8, 50), (64, 20), (128, 10))]
        progress("solver matrix and Laplace done")
    if args.batched_meshes > 0 and not args.no_secondary:
        secondary["batched"] = batched_line(cm, args.batched_meshes)
        progress("batched meshes done")

    line = {
        "metric": "elements/s assembled; time-to-1e-10-residual per load step, plain vs learned",
        "value": value,
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded jittered unit-cube mesh, deformed state)",
        "config": {"workload": f"Mooney-Rivlin {args.n}^3 triquadratic/trilinear, 3x3x3 Gauss, residual + FP64 "
                               "analytic Jacobian assembly into CSR", "elements": E, "dofs": eq.n_rows,
                   "nnz": eq.nnz, "c10": c10, "c01": c01, "jitter": 0.1},
        "roofline": roofline,
        "cpu_baseline": None,
        "secondary": secondary,
    }
    if not args.no_cpu_baseline:
        progress("CPU baseline")
        line["cpu_baseline"] = cpu_baseline(mesh, dofs, c10, c01, args.cpu_seconds)
        line["cpu_baseline"]["host_cpus"] = os.cpu_count()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
