"""Benchmark: FIM cell-steps/s of the three-phase black-oil Newton timestep on one MI355X.

Workload: the 128x128x64 (1,048,576-cell) three-phase deck (seeded lognormal
permeability sigma 1.5, gravity, Leverett-J Pc, five multi-layer wells, 10-day
report steps), Newton tolerance 1e-6 on the scaled residual of every cell family
plus MB <= 1e-7. A "step" is one 10-day report step through the substep
controller; every substep, retry, Newton iteration, assembly and linear solve is
inside the timed region; deck setup is not. The timed simulator starts at time 0
(report step 1, whose first substeps and retries are the transient); the warmup
steps run on a separate simulator of the same deck beforehand.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line. `value` is the mixed-precision throughput; the plain FP64
run of the same steps is reported beside it (`fp64_value`, `mixed_over_fp64`), and
report step 1 in the reference's own ILU ordering (`reference_ordering`), the discrete
adjoint's gradients/s on the 48x48x8 deck (`adjoint`) and a batched two-phase
ensemble's cell-steps/s (`ensemble`).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# per-launch HBM traffic of the pressure SpMV, from the rocprofv3 FETCH_SIZE / WRITE_SIZE
# runs summarised in this file (scaled by the same counters on a copy of known size)
# measured HBM traffic per launch of the kernel the probe times, by kernel
TRAFFIC_FILES = {
    "k_pspmv_dot<float>": os.path.join(ROOT, "profiles", "pressure_spmv_black_oil_128x128x64_traffic.md"),
    "k_rb_a<float>": os.path.join(ROOT, "profiles", "pressure_rb_a_black_oil_128x128x64_traffic.md"),
}
KERNEL_ROLES = {
    "k_pspmv_dot<float>": "pressure SpMV + fused dots, BiCGStab",
    "k_rb_a<float>": ("update of the previous iteration, opening, p and the black half of the pressure ILU on p; "
                      "red-black BiCGStab"),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor() or "unknown"


def solver_options(precision, ordering, args):
    return dict(newton_residual_tol=1e-6, newton_mb_tol=1e-7, min_newton_iteration_index=0,
                precision=precision, ilu_ordering=ordering,
                pressure_max_iterations=args.pressure_max_iterations,
                linear_max_iterations=args.linear_max_iterations, linear_restart=30)


def run_gpu(deck, precision, ordering, args, profile=False, warmup=None, steps=None):
    """Warm up on a separate simulator (`warmup` report steps), then time the first
    `steps` (default `args.steps`) report steps of a fresh one, from time 0."""
    from ressim_amd import ReservoirSimulator

    sched = deck.schedule
    warmup = args.warmup if warmup is None else warmup
    if warmup > 0:
        w = deck.build(ReservoirSimulator)
        w.set_solver_options(**solver_options(precision, ordering, args))
        for k in range(warmup):
            rc = w.step(sched[k])
            log(f"[bench] warmup step {k} rc={rc} substeps={len(w.accepted_dts())}")
            if rc:
                raise RuntimeError("warmup step failed: " + w.last_warning())
        w.close()
    sim = deck.build(ReservoirSimulator)
    sim.set_solver_options(**solver_options(precision, ordering, args))
    sim.debug_initial_state()  # device buffers and the model resident before the timed region
    if profile:
        # time the roofline kernel live over the timed steps: dispatch-timestamped HIP
        # start/stop events on the solver's own stream, on every 8th iteration's launch
        # (timing every launch would add about 8 % to the step time)
        sim.set_kernel_probe(True, every=8)
    s0 = sim.stats()
    per_step = []
    step_s = []
    steps = args.steps if steps is None else steps
    t0 = time.perf_counter()
    for k in range(steps):
        a = sim.stats()
        ts = time.perf_counter()
        rc = sim.step(sched[k])
        step_s.append(time.perf_counter() - ts)
        b = sim.stats()
        per_step.append((b["substeps"] - a["substeps"], b["retries"] - a["retries"]))
        log(f"[bench] timed step {k + 1} rc={rc} t={time.perf_counter() - t0:.2f}s substeps={per_step[-1][0]} "
            f"retries={per_step[-1][1]}")
        if rc:
            raise RuntimeError("timed step failed: " + sim.last_warning())
    elapsed = time.perf_counter() - t0  # rsf_step returns after its stream has drained
    s1 = sim.stats()
    counts = {k: s1[k] - s0[k] for k in ("newton_iterations", "linear_iterations", "substeps", "retries",
                                           "pressure_iterations", "pressure_solves")}
    counts["substeps_per_step"] = [p[0] for p in per_step]
    counts["retries_per_step"] = [p[1] for p in per_step]
    res = dict(elapsed=elapsed, counts=counts, cell_steps_per_s=deck.n_cells * steps / elapsed, step_s=step_s)
    if profile:
        pr = sim.kernel_probe()
        sim.set_kernel_probe(False)
        res["probe"] = pr if pr["launches"] > 0 else None
        res["kernels"] = sim.profile_kernels(args.profile_reps)
    sim.close()
    return res


def adjoint_pass(args):
    """Discrete adjoint on the 48x48x8 deck (50 report steps of 5 d, 36,864 parameters:
    per-cell ln k and porosity): one recorded forward run and one backward pass, mixed
    precision, red-black ILU, the main workload's Newton tolerance. The headline figure
    keeps a checkpoint every 10 substeps and recomputes the states and Jacobians of each
    segment in the backward pass; `every_iterate_kept` is the same run keeping every
    substep's converged iterate (nothing replayed)."""
    from ressim_amd import decks
    from ressim_amd.adjoint import FimAdjoint

    d = decks.black_oil_adjoint(n_steps=50)
    opts = dict(solver_options(1, 1, args))
    out = {}
    for interval in (10, 1):
        fn = FimAdjoint(d, checkpoint_interval=interval, solver_options=opts)
        t0 = time.perf_counter()
        J, dts = fn.forward()
        t1 = time.perf_counter()
        g_lnk, g_phi = fn.backward()
        t2 = time.perf_counter()
        rep = fn.report
        fn.close()
        fwd, bwd = t1 - t0, t2 - t1
        log(f"[bench] adjoint (checkpoint interval {interval}): forward {fwd:.2f}s backward {bwd:.2f}s "
            f"substeps {len(dts)} {rep}")
        out[interval] = {"grads_per_s": round(1.0 / (fwd + bwd), 4), "forward_s": round(fwd, 3),
                         "backward_s": round(bwd, 3), "backward_over_forward": round(bwd / fwd, 3),
                         "objective_sc_m3": J, "substeps": len(dts),
                         "adjoint_linear_iterations": rep["linear_iterations"], "replay_ms": round(rep["replay_ms"], 1),
                         "grad_norm": float(np.sqrt(g_lnk @ g_lnk + g_phi @ g_phi))}
    main = dict(out[10])
    main.update({"workload": d.name, "params": 2 * d.n_cells, "report_steps": len(d.schedule),
                 "checkpoint_interval": 10, "every_iterate_kept": out[1]})
    return main


def ensemble_pass(args):
    """The batched ensemble (include/ressim_ensemble.h): the 256-member recipe of two-phase
    members of uneven size (nx, ny in 32..96, nz in 4..8: 5.6k to 68k cells, 6.6M in all),
    40 report steps of 2 d, every member with its own substep control, all members in one
    packed index space, one captured graph per Newton iteration of every running member;
    FP32 CPR preconditioner (the reference's pressure-stage settings, 50 iterations to 1e-6)
    under FP64 FGMRES and assembly. The tile choices are timed at creation (outside the
    timed region) and reported. A run with a relaxed pressure stage (10 iterations, 1e-2) is
    reported beside it."""
    from ressim_amd import decks
    from ressim_amd.ensemble import BatchedEnsemble, build_members

    mem = [decks.two_phase_member(m) for m in range(args.ensemble_members)]

    def one(**options):
        t0 = time.perf_counter()
        sims = build_members(mem)
        ens = BatchedEnsemble(sims, use_graph=0 if args.ensemble_no_graph else 1, **options)
        setup = time.perf_counter() - t0
        for s in sims:
            s.close()
        rc, el = ens.run(2.0, args.ensemble_reports)
        stats = [ens.member_stats(m) for m in range(len(mem))]
        done = [s for s in stats if s["status"] == "finished"]
        cell_steps = sum(s["n_cells"] for s in done) * args.ensemble_reports
        tuning = ens.tuning()
        opts = ens.options
        ens.close()
        log(f"[bench] ensemble {options or 'defaults'}: {len(mem)} members x {args.ensemble_reports} x 2 d: "
            f"{cell_steps / el:.4e} cell-steps/s in {el:.2f}s, {len(done)} finished")
        return el, setup, stats, done, cell_steps, tuning, opts

    # a relaxed pressure stage (10 iterations, 1e-2) beside the reference's CPR settings
    relaxed = None
    if not args.ensemble_no_relaxed:
        el_r, _, stats_r, done_r, cs_r, _, _ = one(pressure_max_iterations=10, pressure_rel_tol=1e-2)
        relaxed = {"pressure_max_iterations": 10, "pressure_rel_tol": 1e-2, "cell_steps_per_s": round(cs_r / el_r, 1),
                   "elapsed_s": round(el_r, 3), "failed_members": len(mem) - len(done_r),
                   "linear_iterations": sum(s["linear_iterations"] for s in stats_r)}
    el, setup, stats, done, cell_steps, tuning, opts = one()
    return {"workload": f"two_phase_members_0..{len(mem) - 1}", "members": len(mem),
            "cells": sum(d.n_cells for d in mem), "report_steps": args.ensemble_reports, "report_dt_days": 2.0,
            "cell_steps_per_s": round(cell_steps / el, 1), "elapsed_s": round(el, 3), "setup_s": round(setup, 3),
            "failed_members": len(mem) - len(done),
            "substeps": sum(s["substeps"] for s in stats), "retries": sum(s["retries"] for s in stats),
            "newton_iterations": sum(s["newton_iterations"] for s in stats),
            "linear_iterations": sum(s["linear_iterations"] for s in stats),
            "max_member_newton": max(s["newton_iterations"] for s in stats),
            "options": {"use_graph": opts.use_graph, "newton_residual_tol": opts.newton_residual_tol, "newton_mb_tol": opts.newton_mb_tol,
                        "linear_rel_tol": opts.linear_rel_tol, "krylov_dimension": opts.krylov_dimension,
                        "pressure_max_iterations": opts.pressure_max_iterations,
                        "pressure_rel_tol": opts.pressure_rel_tol},
            "autotune": tuning, "relaxed_pressure_stage": relaxed}


def reference_ordering_pass(deck, args, mixed, fp64):
    """The reference's own (natural) ILU ordering on the same deck: report step 1 from time 0
    in both precisions, each ILU solve one launch per sweep, against the red-black runs' own
    report step 1 (the main timed runs above)."""
    out = {"timed_report_steps": "1 from time 0", "ilu_ordering": "natural"}
    for label, prec, rb in (("mixed", 1, mixed), ("fp64", 0, fp64)):
        if rb is None:
            continue
        r = run_gpu(deck, prec, 0, args, warmup=0, steps=1)
        c = r["counts"]
        log(f"[bench] natural ordering {label}: report step 1 {r['elapsed']:.2f}s {c}")
        rb1 = rb["step_s"][0]
        out[label] = {"cell_steps_per_s": round(r["cell_steps_per_s"], 1), "ms": round(r["elapsed"] * 1e3, 1),
                      "newton_iterations": c["newton_iterations"], "linear_iterations": c["linear_iterations"],
                      "pressure_iterations": c["pressure_iterations"], "substeps": c["substeps"],
                      "retries": c["retries"], "redblack_step1_ms": round(rb1 * 1e3, 1),
                      "time_over_redblack": round(r["elapsed"] / rb1, 2)}
    return out


def traffic_per_launch(kernel):
    """Bytes per launch recorded in the committed traffic summary of `kernel`, or None."""
    try:
        for line in open(TRAFFIC_FILES[kernel]):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except (OSError, KeyError):
        pass
    return None


def per_iteration(mixed, fp64):
    """Wall time per Newton iteration and per pressure (BiCGStab) iteration of both runs,
    and their FP64/mixed ratios."""
    out = {}
    for unit, key in (("newton", "newton_iterations"), ("pressure", "pressure_iterations")):
        m = mixed["elapsed"] * 1e3 / max(mixed["counts"][key], 1)
        f = fp64["elapsed"] * 1e3 / max(fp64["counts"][key], 1)
        out[f"mixed_ms_per_{unit}_iteration"] = round(m, 4)
        out[f"fp64_ms_per_{unit}_iteration"] = round(f, 4)
        out[f"fp64_over_mixed_per_{unit}_iteration"] = round(f / m, 3)
    return out


def _threads_available():
    for var in ("OMP_NUM_THREADS",):
        v = os.environ.get(var)
        if v and v.isdigit() and int(v) > 0:
            return int(v)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(args, deck, newton_iterations):
    """The oracle (the CPU FP64 restatement of the reference's algorithm) on the 128x128x64
    deck itself: one Newton iteration at the start of report step 1 (assembly, CPR-FGMRES
    solve with the bench's tolerances, update), timed with 1 thread and with every host
    thread. cell-steps/s is extrapolated: the timed steps' cells x steps over (seconds per
    Newton iteration x the Newton iterations the GPU run took for those steps)."""
    import oracle

    def one(threads):
        used = oracle.set_threads(threads)
        o = deck.build(oracle.OracleSimulator)
        o.set_newton_options(max_iterations=1, min_iteration_index=0, residual_tol=1e-6)
        t0 = time.perf_counter()
        o.run_substep(deck.schedule[0])  # stops after its one Newton iteration
        el = time.perf_counter() - t0
        st = o.stats()
        del o
        return used, el, st

    all_threads = _threads_available()
    res = {}
    for label, th in (("single_thread", 1), ("all_threads", all_threads)):
        used, el, st = one(th)
        log(f"[bench] cpu baseline {label}: {used} threads, one Newton iteration {el:.1f} s "
            f"(FGMRES {st['linear']}, BiCGStab {st['bicgstab']})")
        res[label] = dict(threads=used, s_per_newton_iteration=round(el, 3), fgmres_iterations=st["linear"],
                          bicgstab_iterations=st["bicgstab"],
                          value=deck.n_cells * args.steps / (el * max(newton_iterations, 1)))
    # the reference's own CPU-runnable case, whole: 200x1x1 waterflood, 150 one-day steps with
    # the reference's default options, on the GPU and on the oracle (1 thread)
    from ressim_amd import decks
    from ressim_amd.simulator import ReservoirSimulator

    small = {}
    wf = decks.waterflood_1d()
    oracle.set_threads(1)
    for label, factory in (("gpu", ReservoirSimulator), ("oracle_1_thread", oracle.OracleSimulator)):
        s = wf.build(factory)
        t0 = time.perf_counter()
        for dt in wf.schedule:
            if s.step(dt) != 0:
                raise RuntimeError(f"waterflood_1d failed on {label}")
        small[label + "_s"] = round(time.perf_counter() - t0, 4)
        getattr(s, "close", lambda: None)()
    small["workload"] = "waterflood_200x1x1, 150 x 1 d, default options"
    log(f"[bench] waterflood_200x1x1 150 steps: gpu {small['gpu_s']:.3f} s, oracle {small['oracle_1_thread_s']:.3f} s")
    oracle.set_threads(all_threads)
    best = res["all_threads"]
    host = f"{cpu_model()}, nproc {os.cpu_count()}"
    return dict(value=round(best["value"], 2), unit="cell-steps/s", cores=best["threads"], kind="port",
                sample=(f"oracle (C++ FP64 restatement, OpenMP over independent loops; natural-order ILU and all "
                        f"sums sequential) on the 128x128x64 deck: one Newton iteration at the start of report "
                        f"step 1, {best['s_per_newton_iteration']:.1f} s with {best['threads']} threads, "
                        f"{res['single_thread']['s_per_newton_iteration']:.1f} s with 1 thread; cell-steps/s "
                        f"extrapolated with the {newton_iterations} Newton iterations the GPU mixed run took "
                        f"for the same {args.steps} report steps; host {host}"),
                extrapolated=True, host_cpu=cpu_model(), nproc=os.cpu_count(),
                single_thread=dict(value=round(res["single_thread"]["value"], 2), threads=1,
                                   s_per_newton_iteration=res["single_thread"]["s_per_newton_iteration"]),
                all_threads=dict(value=round(best["value"], 2), threads=best["threads"],
                                 s_per_newton_iteration=best["s_per_newton_iteration"]),
                reference_case=small)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--ordering", choices=["natural", "redblack"], default="redblack")
    # reference defaults: fim/linear/gmres_block_jacobi.rs:25, fim/linear/mod.rs:207-235
    ap.add_argument("--pressure-max-iterations", type=int, default=50)
    ap.add_argument("--linear-max-iterations", type=int, default=20)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-natural", action="store_true")
    ap.add_argument("--ensemble-members", type=int, default=256)
    ap.add_argument("--ensemble-reports", type=int, default=40)
    # launch every tick's kernels directly instead of replaying the captured graph (the
    # rocprofv3 kernel tracer crashes on graph launches)
    ap.add_argument("--ensemble-no-graph", action="store_true")
    ap.add_argument("--ensemble-no-relaxed", action="store_true")
    ap.add_argument("--profile-reps", type=int, default=20)
    ap.add_argument("--deck", default="black_oil_128x128x64")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported in this build "
              "(the multi-GPU decomposition is a design, see the project documentation)", file=sys.stderr)
        sys.exit(2)

    from ressim_amd import decks

    n_total = max(args.warmup, args.steps, 30)
    if args.deck == "black_oil_128x128x64":
        deck = decks.black_oil_1m(n_steps=n_total)
    elif args.deck == "black_oil_64x64x16":
        deck = decks.black_oil_65k(n_steps=n_total)
    elif args.deck == "black_oil_256x256x128":
        deck = decks.black_oil_8m(n_steps=n_total)
    else:
        raise SystemExit("unknown deck " + args.deck)
    ordering = 0 if args.ordering == "natural" else 1
    log(f"[bench] deck {deck.name} cells={deck.n_cells} warmup={args.warmup} steps={args.steps} "
        f"ordering={args.ordering}")

    mixed = run_gpu(deck, 1, ordering, args, profile=True)
    log(f"[bench] mixed: {mixed['cell_steps_per_s']:.4e} cell-steps/s {mixed['counts']}")
    fp64 = None
    if not args.no_fp64:
        fp64 = run_gpu(deck, 0, ordering, args, profile=False, warmup=min(args.warmup, 1))
        log(f"[bench] fp64: {fp64['cell_steps_per_s']:.4e} cell-steps/s {fp64['counts']}")

    k = mixed["kernels"]
    # the BiCGStab kernel with the most GPU time (k_rb_a in red-black ordering with nx even,
    # the pressure SpMV k_pspmv_dot otherwise), timed live over the timed steps
    pr = mixed["probe"]
    roofline = None
    if pr is not None:
        achieved = pr["bytes_per_launch"] / (pr["avg_ms"] * 1e-3) / 1e9
        # the committed traffic summaries measure the 128x128x64 workload only
        tr = traffic_per_launch(pr["kernel"]) if deck.name == "black_oil_128x128x64" else None
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None if tr is None else round(tr),
                    "kernel": f"{pr['kernel']} ({KERNEL_ROLES.get(pr['kernel'], 'BiCGStab')})",
                    "launches": pr["launches"], "avg_us": round(pr["avg_ms"] * 1e3, 3),
                    "algorithmic_bytes": round(pr["bytes_per_launch"])}
        log(f"[bench] probe: {pr['launches']} launches avg {pr['avg_ms'] * 1e3:.2f} us -> {achieved:.0f} GB/s")
    kern = {name: {"ms": round(k[name + "_ms"], 5), "GB/s": round(k[name + "_bytes"] / (k[name + "_ms"] * 1e6), 1)}
            for name in ("spmv", "bilu", "pilu", "pspmv", "assembly", "copy")}

    natural = None if (args.no_natural or ordering == 0) else reference_ordering_pass(deck, args, mixed, fp64)
    adjoint = None if args.no_adjoint else adjoint_pass(args)
    ensemble = ensemble_pass(args) if args.ensemble_members > 0 else None
    cpu = None if args.no_cpu_baseline else cpu_baseline(args, deck, mixed["counts"]["newton_iterations"])
    value = mixed["cell_steps_per_s"]
    out = {
        "metric": "FIM cell-steps/s at 1e-6 Newton tol, mixed vs FP64 GPU; adjoint grads/s",
        "value": round(value, 2),
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["elapsed"] / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 assembly/outer FGMRES, f32 CPR stages",
        "data": "synthetic (seeded GRF permeability, correlation PVT)",
        "config": {"workload": deck.name, "cells": deck.n_cells, "report_step_days": deck.schedule[0],
                   "timed_report_steps": f"1..{args.steps} from time 0 (warmup on a separate simulator)",
                   "newton_tol": 1e-6, "precision": "mixed", "ilu_ordering": args.ordering,
                   "pressure_max_iterations": args.pressure_max_iterations,
                   "linear_max_iterations": args.linear_max_iterations, "counts": mixed["counts"]},
        "fp64_value": None if fp64 is None else round(fp64["cell_steps_per_s"], 2),
        "mixed_over_fp64": None if fp64 is None else round(value / fp64["cell_steps_per_s"], 3),
        "fp64_counts": None if fp64 is None else fp64["counts"],
        # The two runs may take different substep paths (a retry more or less decided by
        # rounding), so the whole-run ratio mixes precision and path. Time per unit of solver
        # work separates them.
        "per_iteration": None if fp64 is None else per_iteration(mixed, fp64),
        "roofline": roofline,
        "kernels": kern,
        "reference_ordering": natural,
        "adjoint": adjoint,
        "ensemble": ensemble,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
