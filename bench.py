#!/usr/bin/env python3
"""Throughput of the FP64 IMEX step on one MI355X.

Workload: the lid-driven cavity with a deformable disk at N = 1024 (ACDI,
seed 0: radius ~ U[0.15, 0.25], centre ~ U[0.4, 0.6]^2, dt = 1e-3*128/N),
solved with the drivers' solver settings under the reference's stopping
rules: the velocity predictor to reltol 1e-8 of its initial residual from
u^n (IterativeSolvers' rule, TimeIntegrationIMEX.jl:602-641, 807), the
projection to 1e-10.  A "step" is one integrate_step! of the whole grid with
the state resident in HBM.  The timed window is steps 20 .. 20+K of the run
from rest (the measurement window of the cavity benchmark): the steps before
it develop the flow, the last W of them are the warm-up.

The JSON line carries:
  roofline     -- the dominant kernel's algorithmic bytes / its mean launch
                  time, measured with HIP events on the library's own stream;
  cpu_baseline -- the CPU oracle (a C restatement of the reference step)
                  timed on bounded samples from the GPU's state after the
                  timed region, same tolerances: its OpenMP build on every
                  usable core (the reported value), and the serial build on
                  one thread ("single_thread").
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

METRIC = "IMEX steps/s at N=1024 (FP64 vs mixed @1e-8), HBM roofline %, adjoint grads/s"
HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s

# Algorithmic HBM passes per step, in fields of N^2 doubles.
def step_bytes(N, vel_iters, proj_iters, direct=True, pcg_passes=17.3, gmres=True):
    """Passes this implementation makes per step: tendencies 107 (momentum 6
    reads per stage plus the running sums of the later stages' velocity
    right-hand sides, 8 + 8 + 4; phase 3 x 4; sigma 7 reads per stage plus the
    running sums of the later stages' phi and sigma, 12 + 16 + 8; the stage
    combinations happen inside these), + 2 under the GMRES stopping rule (the
    stage-1 momentum kernel stores M u^n); per velocity solve 18 (start:
    right-hand side and four history fields read, x r written; q0: r read, q
    written), + 4 under the GMRES rule (u^n and M u^n read for |b - A u^n|),
    + 16 per CG iteration; per projection with the direct transform solve 16
    (first transform pass: U V b_bc read, rhs and the row transform written;
    column and inverse row passes 2 each; the tested update and checks:
    q U V rhs read, U V q_out written), or 17 + 17.3 per PCG iteration when
    the multigrid-PCG solves it."""
    cell = 8.0 * N * N
    per_proj = 16 if direct else 17
    per_solve = 18 + (4 if gmres else 0)
    passes = 107 + (2 if gmres else 0) + sum(per_solve + 16 * k for k in vel_iters) + \
        sum(per_proj + pcg_passes * k for k in proj_iters)
    return passes * cell


# Compulsory passes of a textbook implementation of the same step (the
# accounting the throughput target is quoted in): tendencies 3 x 14, stage
# combinations 80, textbook CG 28 passes per iteration + 6 per solve, textbook
# PCG with a V-cycle 30 per iteration + 18 per projection, report 6.
def textbook_step_bytes(N, vel_iters, proj_iters):
    cell = 8.0 * N * N
    passes = 42 + 80 + sum(6 + 28 * k for k in vel_iters) + sum(18 + 30 * k for k in proj_iters) + 6
    return passes * cell


def profiled_traffic(kernel, N, method):
    """HBM bytes per launch of `kernel` from its committed rocprof summary
    (FETCH_SIZE + WRITE_SIZE, calibrated by a copy kernel), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_cavity{N}_{method}.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def host_cpu():
    """The host the CPU baseline ran on: model name, logical CPUs, the CPUs
    this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = None
    return {"model": model, "logical_cpus": os.cpu_count(), "usable_cpus": usable}


def mixed_filament(fsi, setups, nsteps, window):
    """The wall-supported filament in Kolmogorov flow at N = 1024 (UWCDI, the
    drivers' tolerances) run for `nsteps` steps in FP64 and in mixed
    precision (the predictor by FP64 refinement around an FP32 CG): steps/s
    of both over the steps from `window` on, per-field relative L2 of the
    mixed run against the FP64 run at the end, and the error of one mixed
    step from the FP64 run's final state against the FP64 step."""
    fields = ("U", "V", "phi", "sigma_xx", "sigma_yy", "sigma_xy")

    def run(precision):
        case = setups.kolmogorov_filament(1024, "uwcdi", precision=precision)
        ctx = fsi.build_imex_context(case.metrics, case.bcs, case.models, case.numerics)
        work = fsi.allocate_imex_workspace(ctx)
        fsi.make_velocity_compatible(case.state, ctx, work)
        t, dt = 0.0, case.dt
        for _ in range(window):
            fsi.integrate_step(case.state, t, dt, ctx, work)
            t += dt
        t0 = time.perf_counter()
        its, cyc = [], []
        for _ in range(nsteps - window):
            fsi.integrate_step(case.state, t, dt, ctx, work)
            t += dt
            its.append(work.last_stats.velocity_iterations)
            cyc.append(work.last_stats.velocity_cycles)
        el = time.perf_counter() - t0
        out = {f: getattr(case.state, f).copy() for f in fields}
        print(f"filament {precision}: {1e3 * el / (nsteps - window):.3f} ms/step, velocity iterations "
              f"{np.mean(np.array(its), axis=0)}, cycles {np.mean(np.array(cyc), axis=0)}", file=sys.stderr,
              flush=True)
        return out, (nsteps - window) / el, np.mean(np.array(its), axis=0), np.mean(np.array(cyc), axis=0), \
            (case, ctx, work, t, dt)

    ref, v64, it64, _, (c64, ctx64, w64, t64, dt) = run("fp64")
    mix, vmx, itmx, cymx, _ = run("mixed")

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    vel = np.sqrt(np.sum(ref["U"] ** 2) + np.sum(ref["V"] ** 2))
    run_err = {f: rel(mix[f], ref[f]) for f in fields}
    run_err["UV"] = float(np.sqrt(np.sum((mix["U"] - ref["U"]) ** 2) + np.sum((mix["V"] - ref["V"]) ** 2)) / vel)
    # one step from the same state in both precisions
    mcase = setups.kolmogorov_filament(1024, "uwcdi", precision="mixed")
    mctx = fsi.build_imex_context(mcase.metrics, mcase.bcs, mcase.models, mcase.numerics)
    mwork = fsi.allocate_imex_workspace(mctx)
    st = fsi.FSIState(*(ref[f].copy(order="F") if f in ref else getattr(mcase.state, f).copy(order="F")
                        for f in ("U", "V", "phi", "resource", "sigma_xx", "sigma_yy", "sigma_xy")))
    fsi.integrate_step(st, t64, dt, mctx, mwork)
    fsi.integrate_step(c64.state, t64, dt, ctx64, w64)
    step_err = {f: rel(getattr(st, f), getattr(c64.state, f)) for f in fields}
    print(f"filament mixed vs fp64: after the run {run_err}, one step {step_err}", file=sys.stderr, flush=True)
    return {"workload": "wall-supported filament in Kolmogorov flow, 1024x1024, UWCDI, G = 0.5, mu = 0.01, "
                        f"f_x = 0.5 sin(4 pi y), fixed dt = {dt:.4g}, {nsteps} steps from rest; velocity GMRES "
                        "rule 1e-8, projection 1e-11",
            "value": vmx, "unit": "IMEX steps/s", "fp64_value": v64, "speedup_vs_fp64": vmx / v64,
            "timed_steps": f"{window}..{nsteps}", "rel_l2_vs_fp64_after_run": run_err,
            "rel_l2_vs_fp64_one_step": step_err, "velocity_iterations_fp64": it64.tolist(),
            "velocity_inner_iterations_mixed": itmx.tolist(), "velocity_cycles_mixed": cymx.tolist(),
            "scheme": "velocity predictor: FP64 iterative refinement around an FP32 CG (FP32 r, p, q; FP64 x, "
                      "residual checks and dot products; the same stopping rule); projection V-cycle in FP32; "
                      "state, tendencies, stage sums, the direct projection solve and every reduction FP64",
            "_states": None}


def adjoint_leg(fsi, setups, N, nsteps, snapshots, fd_check=True):
    """The differentiation benchmark: lid-driven cavity, ACDI, FP64, tight
    solves, `nsteps` steps of dt = 1e-3 * 128 / N (the largest step of the
    cavity benchmarks that keeps ACDI's explicit diffusion stable at N = 256:
    dt = 1e-3 there gives 8 eps Gamma dt / h^2 = 3.1, beyond the RK3
    stability interval, and the run fails with a non-SPD sigma);
    J = |x_c(phi(T)) - x*|^2 with x*
    drawn from seed 5; dJ/d(shear modulus, disk radius, disk centre) by the
    device adjoint with binomial checkpointing (`snapshots` states), and once
    with every step's tape kept (no recomputation: 19 fields per step of HBM);
    the plain forward run timed beside it; central
    differences (h = 1e-6) of the device's forward runs as the check."""
    from eulerianfsi_amd import adjoint as A

    rng = setups.SplitMix64(5)
    target = (rng.uniform(0.4, 0.6), rng.uniform(0.4, 0.6))
    P0 = (0.1, 0.2, 0.6, 0.5)
    dt = 1e-3 * 128 / N

    def case_of(p):
        G, r, cx, cy = p
        return setups.lid_driven_cavity(N, "acdi", radius=r, center=(cx, cy), shear_modulus=G, dt=dt,
                                        velocity_method="direct", projection_rtol=1e-13)

    def forward_J(p):
        c = case_of(p)
        ctx = fsi.build_imex_context(c.metrics, c.bcs, c.models, c.numerics)
        work = fsi.allocate_imex_workspace(ctx)
        t0 = time.perf_counter()
        for k in range(nsteps):
            fsi.integrate_step(c.state, k * c.dt, c.dt, ctx, work)
        _ = c.state.phi
        el = time.perf_counter() - t0
        J, _ = A.centroid_objective(c.state, ctx, target)
        return J, el

    J0, t_fwd = forward_J(P0)
    print(f"adjoint leg: forward run {nsteps} steps at N={N}: {t_fwd:.3f} s, J = {J0:.6e}", file=sys.stderr,
          flush=True)
    runs = {}
    for snaps in (snapshots, 0):
        out = A.disk_gradient(case_of(P0), nsteps, target, snapshots=snaps)
        st = out["stats"]
        el = out["forward_s"] + out["backward_s"]
        runs[snaps] = dict(grad=out["grad"].tolist(), seconds=el, forward_s=out["forward_s"],
                           backward_s=out["backward_s"], grads_per_s=1.0 / el, cost_vs_forward=el / t_fwd,
                           recomputed_steps=st.advances, taped_steps=st.recordings, snapshots_held=st.snapshots_used,
                           repetitions=st.repetitions, adjoint_device_bytes=st.device_bytes, J=out["J"])
        print(f"adjoint leg: {snaps or 'all tapes'} snapshots: {el:.3f} s ({el / t_fwd:.2f}x the forward run), gradient "
              f"{out['grad']}, recomputed {st.advances}, {st.device_bytes / 2**20:.0f} MiB", file=sys.stderr,
              flush=True)
        del out
    res = {"workload": f"lid-driven cavity {N}x{N}, ACDI, FP64, G = 0.1, disk r = 0.2 at (0.6, 0.5), dt = {dt:g}, "
                       f"{nsteps} steps (T = {nsteps * dt:g}); J = |x_c(phi(T)) - x*|^2, x* = "
                       f"({target[0]:.4f}, {target[1]:.4f}) from seed 5; dJ/d(G, r, cx, cy); velocity :direct "
                       f"(CG to 1e-14), projection rtol 1e-13",
           "J": J0, "forward_run_s": t_fwd, "binomial": runs[snapshots], "all_tapes_kept": runs[0],
           "snapshots": snapshots}
    if fd_check:
        h = 1e-6
        g = np.array(runs[snapshots]["grad"])
        fd = []
        for k in range(4):
            pp, pm = list(P0), list(P0)
            pp[k] += h
            pm[k] -= h
            fd.append((forward_J(pp)[0] - forward_J(pm)[0]) / (2 * h))
            print(f"adjoint leg: central difference {k}: {fd[-1]:.10e} vs adjoint {g[k]:.10e}", file=sys.stderr,
                  flush=True)
        fd = np.array(fd)
        res["central_differences"] = fd.tolist()
        res["rel_diff_vs_central_differences"] = (np.abs(fd - g) / np.abs(g)).tolist()
        res["fd_step"] = h
    return res


def growth_leg(fsi, setups, N, nsteps, warmup):
    """The growth and resource step (growing disk, setups.growing_disk) under
    the reference's stopping rules: velocity and resource GMRES rule 1e-8,
    projection 1e-10.  Every solve is polled (no planned steps); each
    integrate_step returns after its stream has finished (the Status read-back),
    so the host clock brackets completed work."""
    numerics = fsi.IMEXNumerics(projection_reltol=1e-10, linear_solvers=fsi.LinearSolverOptions(
        velocity_method="gmres", resource_method="gmres", reltol=1e-8))
    grid, metrics, bcs, models, numerics, state = setups.growing_disk(N, numerics=numerics)
    ctx = fsi.build_imex_context(metrics, bcs, models, numerics)
    work = fsi.allocate_imex_workspace(ctx)
    dt = 1e-3 * 128 / N
    fsi.make_velocity_compatible(state, ctx, work)
    t = 0.0
    for _ in range(warmup):
        fsi.integrate_step(state, t, dt, ctx, work)
        t += dt
    t0 = time.perf_counter()
    rep = None
    for _ in range(nsteps):
        rep = fsi.integrate_step(state, t, dt, ctx, work)
        t += dt
    el = time.perf_counter() - t0
    return {"workload": f"elastic ACDI disk growing on a diffused, consumed resource, {N}x{N}, FP64, dt = {dt:.3g}, "
                        f"steps {warmup}..{warmup + nsteps} from rest; SaturatingResourceGrowth(0.5, 0.5), "
                        "DirectionalGrowthTensor(1, 2), HarmonicPhaseDiffusivity(0.05, 0.01), "
                        "SaturatingResourceConsumption(0.2, 0.5); velocity and resource GMRES rule 1e-8, "
                        "projection 1e-10 (multigrid PCG: the opened top has no transform basis)",
            "value": nsteps / el, "unit": "IMEX steps/s", "ms_per_step": 1e3 * el / nsteps,
            "cell_updates_per_s": N * N * nsteps / el,
            "maximum_growth_trace_error": rep.maximum_growth_trace_error, "minimum_resource": rep.minimum_resource,
            "maximum_projection_constraint_linf": rep.maximum_projection_constraint_linf}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--window-start", type=int, default=20,
                    help="step of the run from rest at which the timed window starts (>= warmup)")
    ap.add_argument("--N", type=int, default=1024)
    ap.add_argument("--method", default="acdi")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=2, help="oracle steps in the CPU baseline sample (0: skip)")
    ap.add_argument("--cpu-omp-steps", type=int, default=5,
                    help="steps of the OpenMP build of the oracle on all usable cores (0: skip)")
    ap.add_argument("--kernel-reps", type=int, default=50)
    ap.add_argument("--only-kernel", default=None, help="time one kernel standalone and exit (profiling runs)")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision run")
    ap.add_argument("--mixed-steps", type=int, default=2000, help="steps of the filament run in each precision")
    ap.add_argument("--mixed-window", type=int, default=20, help="untimed first steps of the filament runs")
    ap.add_argument("--ensemble-steps", type=int, default=10,
                    help="timed steps of the 64-cavity ensemble, stepped one by one and batched (0: skip)")
    ap.add_argument("--ensemble-warmup", type=int, default=4)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"], help="precision of the main run")
    ap.add_argument("--adjoint-steps", type=int, default=2000, help="steps of the differentiation run (0: skip)")
    ap.add_argument("--adjoint-N", type=int, default=256)
    ap.add_argument("--adjoint-snapshots", type=int, default=45)
    ap.add_argument("--no-adjoint-fd", action="store_true", help="skip the central-difference check")
    ap.add_argument("--growth-steps", type=int, default=10, help="timed steps of the growth leg (0: skip)")
    ap.add_argument("--growth-N", type=int, default=1024)
    ap.add_argument("--only-growth", action="store_true",
                    help="run the growth leg alone and print its JSON (profiling runs)")
    ap.add_argument("--only-adjoint", action="store_true",
                    help="run the differentiation leg alone and print its JSON (profiling runs)")
    args = ap.parse_args(argv)

    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build's step runs on a single MI355X (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        return 2
    if args.only_growth:
        import eulerianfsi_amd as fsi
        from eulerianfsi_amd import setups
        print(json.dumps(growth_leg(fsi, setups, args.growth_N, args.growth_steps, 3)))
        return 0
    if args.only_adjoint:
        import eulerianfsi_amd as fsi
        from eulerianfsi_amd import setups
        print(json.dumps(adjoint_leg(fsi, setups, args.adjoint_N, args.adjoint_steps, args.adjoint_snapshots,
                                     fd_check=not args.no_adjoint_fd)))
        return 0

    import eulerianfsi_amd as fsi
    from eulerianfsi_amd import _abi, setups

    case = setups.seeded_cavity(args.N, args.method, args.seed, velocity_rtol=1e-8, projection_rtol=1e-10,
                                precision=args.precision)
    ctx = fsi.build_imex_context(case.metrics, case.bcs, case.models, case.numerics)
    work = fsi.allocate_imex_workspace(ctx)
    state = case.state
    dt = case.dt
    t = 0.0
    fsi.make_velocity_compatible(state, ctx, work)
    start = max(args.window_start, args.warmup)
    for k in range(start):
        fsi.integrate_step(state, t, dt, ctx, work)
        t += dt
        if k >= start - args.warmup:
            print(f"warmup step {k} t={t:.5f} iters={work.last_stats}", file=sys.stderr, flush=True)
        else:
            print(f"development step {k} t={t:.5f} iters={work.last_stats.velocity_iterations}", file=sys.stderr,
                  flush=True)

    def bench_kernel(name, reps):
        import ctypes
        ms, nbytes = ctypes.c_double(), ctypes.c_double()
        st = ctx.lib.fsi_benchmark_kernel(ctx.handle, state._device(ctx), work.handle, name.encode(), reps,
                                          ctypes.byref(ms), ctypes.byref(nbytes))
        _abi.check(ctx.lib, ctx.handle, st, f"benchmark {name}")
        return ms.value, nbytes.value

    if args.only_kernel:
        ms, nb = bench_kernel(args.only_kernel, args.kernel_reps)
        print(json.dumps({"kernel": args.only_kernel, "avg_ms": ms, "bytes": nb, "GBps": nb / ms / 1e6}))
        return 0

    vel_tot, proj_tot, vres, vref, planned = [], [], [], [], []
    t0 = time.perf_counter()
    for k in range(args.steps):
        rep = fsi.integrate_step(state, t, dt, ctx, work)
        t += dt
        st = work.last_stats
        vel_tot.append(st.velocity_iterations)
        proj_tot.append(st.projection_iterations)
        vres.append(st.velocity_residual)
        vref.append(st.velocity_reference_residual)
        planned.append(st.planned)
    elapsed = time.perf_counter() - t0
    # the predictor solves against the reference's rule: final |r| / |b - A u^n| per stage
    achieved = (np.array(vres) / np.maximum(np.array(vref), 1e-300)).max(axis=0)
    print(f"velocity solves: max |r|/|b - A u^n| per stage {achieved}, |b - A u^n|/|b| "
          f"{np.mean(np.array(vref), axis=0)}", file=sys.stderr, flush=True)
    steps_per_s = args.steps / elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    vel_mean = np.mean(np.array(vel_tot), axis=0)
    proj_mean = np.mean(np.array(proj_tot), axis=0)
    final = {f: getattr(state, f).copy() for f in ("U", "V", "phi", "sigma_xx", "sigma_yy", "sigma_xy")}
    print(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step, vel iters {vel_mean}, proj iters {proj_mean}, "
          f"report {rep}", file=sys.stderr, flush=True)

    # per-kernel measurement: the candidates for the dominant kernel
    kernels = {}
    for name in ("vel_iter", "spec_solve", "pcg_dir", "pcg_upd", "mg_down0", "mg_up0", "mg_coarse", "pcg_iter",
                 "momentum", "cells", "copy"):
        try:
            ms, nb = bench_kernel(name, args.kernel_reps)
        except fsi.ArgumentError:  # the direct solve is off (FSI_SPECTRAL=0, or a grid it does not cover)
            continue
        kernels[name] = {"avg_ms": ms, "bytes": nb, "GBps": nb / ms / 1e6}
        print(f"kernel {name}: {ms * 1e3:.1f} us, {nb / ms / 1e6:.0f} GB/s", file=sys.stderr, flush=True)
    # launches per step of each kernel (one velocity-CG kernel per iteration,
    # one finest down/up pair and one dir/upd pair per PCG iteration, one
    # direct solve per projection)
    nv = float(np.sum(vel_mean))
    npj = float(np.sum(proj_mean))
    launches = {"vel_iter": nv + 3, "pcg_dir": npj, "pcg_upd": npj, "mg_down0": npj, "mg_up0": npj,
                "mg_coarse": npj, "momentum": 3, "cells": 3}
    direct = "spec_solve" in kernels
    if direct:
        launches["spec_solve"] = 3
    share = {k: launches[k] * kernels[k]["avg_ms"] for k in launches}
    dominant = max(share, key=share.get)
    dk = kernels[dominant]
    roofline = {"kernel": dominant, "bound": "hbm", "achieved": dk["GBps"], "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": dk["GBps"] * 1e9 / HBM_PEAK, "traffic": profiled_traffic(dominant, args.N, args.method),
                "bytes_per_launch": dk["bytes"], "avg_launch_us": dk["avg_ms"] * 1e3,
                "share_of_step": share[dominant] / ms_per_step}

    b_step = step_bytes(args.N, vel_mean, proj_mean, direct)
    b_text = textbook_step_bytes(args.N, vel_mean, proj_mean)
    whole_step = {"algorithmic_GB": b_step / 1e9, "achieved_GBps": b_step / (ms_per_step * 1e-3) / 1e9,
                  "frac_of_peak": b_step / (ms_per_step * 1e-3) / HBM_PEAK,
                  "textbook_GB": b_text / 1e9, "textbook_frac_of_peak": b_text / (ms_per_step * 1e-3) / HBM_PEAK}

    # mixed precision: the same run (same seed, W + K steps) with the FP32
    # V-cycle preconditioner; error of every field against this FP64 run
    mixed = None
    if not args.no_mixed and args.precision == "fp64":
        mixed = mixed_filament(fsi, setups, args.mixed_steps, args.mixed_window)
        del mixed["_states"]

    # batched ensemble: 64 cavities (N in {128, 192, 256, 384}, G, disk drawn
    # from a seed), the same steps once one instance after another
    # (integrate_step) and once batched (ensemble_step, every instance's step
    # in flight together); aggregate cell updates per second
    ensemble = None
    if args.ensemble_steps > 0:
        ecases = setups.ensemble_cases(velocity_rtol=1e-8, projection_rtol=1e-10)
        ectx, ework, ets = [], [], []
        for c in ecases:
            cx = fsi.build_imex_context(c.metrics, c.bcs, c.models, c.numerics)
            wk = fsi.allocate_imex_workspace(cx)
            fsi.make_velocity_compatible(c.state, cx, wk)
            ectx.append(cx)
            ework.append(wk)
            ets.append(0.0)
        states, dts = [c.state for c in ecases], [c.dt for c in ecases]
        cells = sum(c.params["N"] ** 2 for c in ecases)

        def batched_step():
            fsi.ensemble_step(states, ets, dts, ectx, ework)
            for k in range(len(ets)):
                ets[k] += dts[k]

        for _ in range(args.ensemble_warmup):  # the first steps poll and record iteration counts
            batched_step()
        print(f"ensemble: {len(ecases)} instances, {cells / 1e6:.2f} M cells, warm", file=sys.stderr, flush=True)
        s0 = time.perf_counter()
        for _ in range(args.ensemble_steps):
            for k in range(len(ecases)):
                fsi.integrate_step(states[k], ets[k], dts[k], ectx[k], ework[k])
                ets[k] += dts[k]
        t_seq = time.perf_counter() - s0
        s0 = time.perf_counter()
        for _ in range(args.ensemble_steps):
            batched_step()
        t_ens = time.perf_counter() - s0
        ensemble = {"workload": "64 lid-driven cavities, N ~ U{128, 192, 256, 384}, G log-uniform [0.05, 0.5], "
                                "disk r ~ U[0.15, 0.25], centre ~ U[0.4, 0.6]^2, ACDI, FP64, dt = 1e-3*128/N, seed 4",
                    "instances": len(ecases), "cells": cells, "steps": args.ensemble_steps,
                    "cell_updates_per_s_batched": cells * args.ensemble_steps / t_ens,
                    "cell_updates_per_s_one_by_one": cells * args.ensemble_steps / t_seq,
                    "speedup": t_seq / t_ens}
        print(f"ensemble: one by one {t_seq:.3f} s, batched {t_ens:.3f} s, speedup {t_seq / t_ens:.2f}",
              file=sys.stderr, flush=True)
        del ework, ectx, states

    adj = None
    if args.adjoint_steps > 0 and args.precision == "fp64":
        adj = adjoint_leg(fsi, setups, args.adjoint_N, args.adjoint_steps, args.adjoint_snapshots,
                          fd_check=not args.no_adjoint_fd)

    cpu = None
    if args.cpu_steps > 0 or args.cpu_omp_steps > 0:
        from oracle.oracle import OMP_LIB_PATH, Oracle, state_dict

        def oracle_leg(nsteps, lib_path, label):
            o = Oracle(case.metrics, case.bcs, case.models, case.numerics, lib_path=lib_path)
            s = state_dict(state)
            o.correction[:] = work.correction
            c0 = time.perf_counter()
            for k in range(nsteps):
                o.integrate_step(s, t + k * dt, dt)
                print(f"cpu oracle ({label}) step {k} done", file=sys.stderr, flush=True)
            return nsteps / (time.perf_counter() - c0)

        sample = (f"step(s) of the same N={args.N} {args.method} cavity from the GPU state after the timed region, "
                  f"same stopping rules")
        serial = None
        if args.cpu_steps > 0:
            serial = {"value": oracle_leg(args.cpu_steps, None, "1 thread"), "unit": "IMEX steps/s", "cores": 1,
                      "kind": "port", "sample": f"{args.cpu_steps} {sample}, 1 thread"}
        cpu = serial
        if args.cpu_omp_steps > 0:
            v = oracle_leg(args.cpu_omp_steps, OMP_LIB_PATH, "OpenMP")
            import ctypes
            threads = ctypes.CDLL("libgomp.so.1").omp_get_max_threads()
            cpu = {"value": v, "unit": "IMEX steps/s", "cores": threads, "kind": "port",
                   "sample": f"{args.cpu_omp_steps} {sample}, the oracle's OpenMP build (cell loops on {threads} "
                             f"threads, red-black multigrid smoother)",
                   "single_thread": serial}
        cpu["host"] = host_cpu()

    growth = growth_leg(fsi, setups, args.growth_N, args.growth_steps, 3) if args.growth_steps > 0 else None

    out = {
        "metric": METRIC, "value": steps_per_s, "unit": "IMEX steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded cavity initial condition)",
        "config": {"workload": f"lid-driven cavity with deformable disk, N={args.N}x{args.N}, {args.method.upper()}, "
                               f"FP64, dt={dt:.4g}, seed {args.seed}, steps {start}..{start + args.steps} of the run "
                               f"from rest; velocity predictor to 1e-8 of its initial residual from u^n (the "
                               f"reference's GMRES reltol rule), projection rtol 1e-10",
                   "velocity_residual_over_reference": achieved.tolist(),
                   "N": args.N, "method": args.method, "cell_updates_per_s": steps_per_s * args.N ** 2,
                   "velocity_iterations": vel_mean.tolist(), "projection_iterations": proj_mean.tolist(),
                   "planned_steps": f"{sum(planned)}/{len(planned)}",
                   "precision": args.precision, "mixed_precision": mixed, "ensemble": ensemble,
                   "adjoint_grads_per_s": (max(adj["binomial"]["grads_per_s"], adj["all_tapes_kept"]["grads_per_s"])
                                           if adj else None),
                   "adjoint": adj, "growth": growth},
        "roofline": roofline, "whole_step": whole_step, "kernels": kernels, "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
