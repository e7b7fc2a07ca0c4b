"""Training and evaluation of the learned warm start.

Data come from the plain method itself: seeded Mooney-Rivlin uniaxial solves
drawn from the batched-meshes distribution (n ~ U{n_min..n_max},
c10 ~ U[1,5], c01 ~ U[0,10], stretch ~ U[5%,20%], 0.1h jitter), each load
step solved by plain FP64 Newton to a 1e-10 relative residual.  Every load
step gives one sample per quadratic node: features (X, z - X, load increment,
c10, c01) at the start of the step and targets (z_end - z_start, p_end -
p_start interpolated to the node) / increment.

Evaluation solves the same load steps twice, plain (previous solution plus
the boundary increment) and learned (plus the MLP increment), both to
1e-10 times the plain start's residual norm, so both answers meet the same
FP64 residual, and reports Newton iterations and wall time per step.

Usage: python -m cm_amd.train_warm_start --solves 120 --out cm_amd/data/warm_start_mlp.npz
"""
import argparse
import json
import os
import time

import numpy as np

from . import _lib
from .equations_set import MooneyRivlinEquationsSet, initialise, newton_options, synchronise
from .mesh import RegularMesh
from .warm_start import WarmStartMLP, step_features, step_targets

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "data", "warm_start_mlp.npz")


def draw_case(rng, n_min=4, n_max=16):
    return dict(n=int(rng.integers(n_min, n_max + 1)), c10=float(rng.uniform(1, 5)), c01=float(rng.uniform(0, 10)),
                stretch=float(rng.uniform(0.05, 0.20)), seed=int(rng.integers(0, 2 ** 31)))


# Forcing of the plain method: every Newton step's linear system solved to the
# reference's default GMRES relative tolerance (solver_routines.f90:10983; the
# reference sets no Eisenstat-Walker forcing).  The learned form uses
# Eisenstat-Walker forcing capped by ||F||/||F_ref|| (krylov rtol -1).
PLAIN_KRYLOV_RTOL = 1e-5


def solver_options(precision=_lib.OC_FP64, restart=100, reference_norm=0.0, krylov_rtol=-1.0, **extra):
    opts = dict(relative_tolerance=1e-10, absolute_tolerance=0.0, reference_norm=reference_norm,
                krylov_restart=restart, krylov_relative_tolerance=krylov_rtol, jacobian_precision=precision,
                preconditioner=_lib.OC_PRECONDITIONER_MULTIGRID, maximum_iterations=100,
                krylov_maximum_iterations=3000,
                # judged by the residual alone: no step-size test, no evaluation limit
                solution_tolerance=0.0, maximum_function_evaluations=-1)
    opts.update(extra)
    return newton_options(**opts)


def make_problem(case, jitter=0.1):
    mesh = RegularMesh.generate(2, case["n"], jitter=jitter, seed=case["seed"])
    eq = MooneyRivlinEquationsSet(mesh, case["c10"], case["c01"])
    eq.set_dofs(eq.undeformed_dofs())
    eq.fixed_dofs = eq.uniaxial_boundary_conditions(1.0 + case["stretch"])
    eq.set_multigrid()
    return mesh, eq


def collect(case, steps):
    """Plain FP64 solve of one case; returns (features, targets) of every step."""
    mesh, eq = make_problem(case)
    xs, ts = [], []
    dload = case["stretch"] / steps
    for s in range(1, steps + 1):
        d0 = eq.get_dofs()
        eq.load_increment_apply(s, steps)
        eq.newton_solve(solver_options())
        d1 = eq.get_dofs()
        xs.append(step_features(mesh, d0, dload, case["c10"], case["c01"]))
        ts.append(step_targets(mesh, d0, d1, dload, case["c10"], case["c01"]))
    eq.destroy()
    return np.concatenate(xs), np.concatenate(ts)


def generate(n_solves, steps, seed, n_min, n_max, log=print):
    rng = np.random.default_rng(seed)
    X, T = [], []
    t0 = time.time()
    for k in range(n_solves):
        case = draw_case(rng, n_min, n_max)
        x, t = collect(case, steps)
        X.append(x)
        T.append(t)
        if (k + 1) % 10 == 0:
            log(f"  data: {k + 1}/{n_solves} solves, {sum(len(a) for a in X)} samples, {time.time() - t0:.1f} s")
    return np.concatenate(X), np.concatenate(T)


def generate_batched(n_solves, steps, seed, n_min, n_max, net=None, batch_size=256, samples_per_step=2000,
                     log=print):
    """Training data from batches of meshes solved together (one equations
    set per batch): each load step starts from the network's guess when a
    network is given (the answers do not depend on the start: every accepted
    step meets 1e-10 x the plain start's residual, FP64) and falls back to the
    plain start with FP64 Jacobians when that solve does not converge.  Up to
    `samples_per_step` random nodes per mesh and step are kept."""
    from .batch_workload import build
    rng = np.random.default_rng(seed)
    pick = np.random.default_rng(seed + 1)
    cases = [draw_case(rng, n_min, n_max) for _ in range(n_solves)]
    X, T = [], []
    t0 = time.time()
    fallbacks = 0
    for b0 in range(0, n_solves, batch_size):
        cs = cases[b0:b0 + batch_size]
        meshes, batch, _ = build(cs)
        batch.set_multigrid()
        dload = np.array([c["stretch"] / steps for c in cs])
        for s in range(1, steps + 1):
            d0 = batch.get_dofs()
            batch.load_increment_apply(s, steps)
            refs = batch.mesh_residual_norms()
            ok = False
            if net is not None:
                batch.set_dofs(d0)
                net.warm_start(batch, dload)
                batch.load_increment_apply(s, steps)
                st, _ = batch.newton_solve(solver_options(_lib.OC_FP32), reference_norms=refs, raise_on_failure=False)
                ok = bool(st["converged"])
            if not ok:
                if net is not None:
                    fallbacks += 1
                    batch.set_dofs(d0)
                    batch.load_increment_apply(s, steps)
                batch.newton_solve(solver_options(_lib.OC_FP64), reference_norms=refs)
            d1 = batch.get_dofs()
            for m, (mesh, c) in enumerate(zip(meshes, cs)):
                idx = batch.mesh_dofs(m)
                x = step_features(mesh, d0[idx], dload[m], c["c10"], c["c01"])
                t = step_targets(mesh, d0[idx], d1[idx], dload[m], c["c10"], c["c01"])
                k = pick.choice(len(x), size=min(len(x), samples_per_step), replace=False)
                X.append(x[k])
                T.append(t[k])
        batch.destroy()
        log(f"  data: {min(b0 + batch_size, n_solves)}/{n_solves} solves, {sum(len(a) for a in X)} samples, "
            f"{fallbacks} plain fallbacks, {time.time() - t0:.1f} s")
    return np.concatenate(X), np.concatenate(T)


def train(net, X, T, steps, batch=4096, lr=1e-3, seed=1, log=print):
    net.set_dataset(X, T)
    loss = None
    done = 0
    for phase_steps, phase_lr in ((steps // 2, lr), (steps // 4, lr * 0.3), (steps - steps // 2 - steps // 4, lr * 0.1)):
        if phase_steps <= 0:
            continue
        loss = net.train(phase_steps, batch=batch, learning_rate=phase_lr, seed=seed + done)
        done += phase_steps
        log(f"  train: {done}/{steps} steps, lr {phase_lr:.1e}, minibatch loss {loss:.4e}")
    return loss


def solve_steps(case, steps, net=None, precision=_lib.OC_FP64, max_steps=None, restart=100, log=None,
                krylov_rtol=None):
    """Solves the load steps; with `net`, each step starts from the MLP guess.
    Krylov forcing: Eisenstat-Walker with `net`, else PLAIN_KRYLOV_RTOL, unless
    krylov_rtol is given.  Returns per-step dicts (iterations, seconds, residual norms)."""
    if krylov_rtol is None:
        krylov_rtol = -1.0 if net is not None else PLAIN_KRYLOV_RTOL
    mesh, eq = make_problem(case)
    dload = case["stretch"] / steps
    out = []
    for s in range(1, (max_steps or steps) + 1):
        d0 = eq.get_dofs()
        # the plain start's residual norm: the common reference of both methods
        eq.load_increment_apply(s, steps)
        ref = eq.residual_norm()
        t0 = time.perf_counter()
        if net is not None:
            eq.set_dofs(d0)
            net.warm_start(eq, dload)
            eq.load_increment_apply(s, steps)
        st = eq.newton_solve(solver_options(precision, restart, reference_norm=ref, krylov_rtol=krylov_rtol),
                             raise_on_failure=False)
        synchronise()
        wall = time.perf_counter() - t0
        out.append(dict(step=s, converged=bool(st["converged"]), newton=st["newton_iterations"],
                        krylov=st["krylov_iterations"], seconds=wall, start_norm=st["initial_residual_norm"],
                        reference_norm=ref, final_norm=st["final_residual_norm"]))
        if log is not None:
            log(f"    {'learned' if net is not None else 'plain'} step {s}: {out[-1]}")
        if not st["converged"]:
            break
    eq.destroy()
    return out


def summarise(rows):
    ok = [r for r in rows if r["converged"]]
    return dict(steps=len(rows), converged=len(ok), newton_mean=float(np.mean([r["newton"] for r in rows])),
                seconds_mean=float(np.mean([r["seconds"] for r in rows])),
                krylov_mean=float(np.mean([r["krylov"] for r in rows])),
                start_over_reference=float(np.mean([r["start_norm"] / r["reference_norm"] for r in rows])))


def main():
    ap = argparse.ArgumentParser(description="train and evaluate the learned warm start")
    ap.add_argument("--solves", type=int, default=120)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--n-min", type=int, default=4)
    ap.add_argument("--n-max", type=int, default=8)
    ap.add_argument("--train-steps", type=int, default=4000)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--heldout", type=int, default=6)
    ap.add_argument("--big", type=int, default=0, help="also evaluate the n^3 case with 20%% stretch in 10 steps")
    ap.add_argument("--big-steps", type=int, default=10)
    ap.add_argument("--report", default=None)
    ap.add_argument("--krylov-rtol", type=float, default=PLAIN_KRYLOV_RTOL,
                    help="GMRES forcing of both methods in the evaluation (<= 0: Eisenstat-Walker)")
    ap.add_argument("--eval-only", action="store_true", help="evaluate the weights at --out without training")
    ap.add_argument("--batched", action="store_true",
                    help="generate the data with batched solves (256 meshes per equations set)")
    ap.add_argument("--bootstrap", default=None,
                    help="weights used as the warm start while generating batched data")
    args = ap.parse_args()
    initialise(0)
    log = lambda m: print(m, flush=True)
    if args.eval_only:
        net = WarmStartMLP.load(args.out)
        report = dict(training=None, heldout=[])
    else:
        t0 = time.time()
        if args.batched:
            boot = WarmStartMLP.load(args.bootstrap) if args.bootstrap else None
            X, T = generate_batched(args.solves, args.steps, args.seed, args.n_min, args.n_max, net=boot, log=log)
        else:
            X, T = generate(args.solves, args.steps, args.seed, args.n_min, args.n_max, log)
        log(f"data: {len(X)} samples from {args.solves} solves in {time.time() - t0:.1f} s")
        net = WarmStartMLP(hidden=128, layers=4, seed=args.seed)
        t1 = time.time()
        loss = train(net, X, T, args.train_steps, log=log)
        log(f"trained in {time.time() - t1:.1f} s, loss {loss:.4e}")
        os.makedirs(os.path.dirname(args.out), exist_ok=True)
        net.save(args.out)
        report = dict(training=dict(solves=args.solves, samples=int(len(X)), steps=args.train_steps, loss=loss,
                                    n_range=[args.n_min, args.n_max]), heldout=[])
    rng = np.random.default_rng(args.seed + 7919)
    for k in range(args.heldout):
        case = draw_case(rng, 4, 16)
        plain = summarise(solve_steps(case, args.steps, krylov_rtol=args.krylov_rtol))
        learned = summarise(solve_steps(case, args.steps, net=net, precision=_lib.OC_FP32, krylov_rtol=args.krylov_rtol))
        report["heldout"].append(dict(case=case, plain_fp64=plain, learned_fp32=learned))
        log(f"heldout {case}: plain {plain} | learned {learned}")
        if args.report:
            with open(args.report, "w") as f:
                json.dump(report, f, indent=1)
    if args.big:
        case = dict(n=args.big, c10=2.0, c01=6.0, stretch=0.20, seed=2024)
        learned = solve_steps(case, args.big_steps, net=net, precision=_lib.OC_FP32, log=log,
                              krylov_rtol=args.krylov_rtol)
        plain = solve_steps(case, args.big_steps, log=log, krylov_rtol=args.krylov_rtol)
        report["big"] = dict(case=case, plain_fp64=summarise(plain), learned_fp32=summarise(learned),
                             plain_rows=plain, learned_rows=learned)
        log(f"big: plain {summarise(plain)} | learned {summarise(learned)}")
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)


if __name__ == "__main__":
    main()
