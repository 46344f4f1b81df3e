#!/usr/bin/env python3
"""Benchmark: cell-updates/s per RK3 step of the 256×256×128 FP32 free-convection LES
(WENO5 + AnisotropicMinimumDissipation + BuoyancyTracer + FPlane + surface buoyancy flux,
FFT pressure solver, Δt = 1 s) on one MI355X, through the C ABI.

Protocol (benchmarking/src/utils.jl:84-101): `--warmup` untimed steps, device sync, `--steps`
timed steps, device sync.  Inputs are resident in HBM before the timed region.  Per-phase device
times are HIP events recorded on the model's stream inside the timed region; the roofline entry
is the dominant kernel's.  The CPU baseline is the FP64 oracle (a CPU restatement of the
reference algorithm) timed on this host: the full size on all physical cores, and one thread on
a bounded sample, with the CPU model.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "oceananigans.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0       # MI355X HBM3E
N_TRACERS = 1


def algorithmic_bytes_per_cell_step(n_prognostic, word):
    """Compulsory traffic per cell per RK3 step, W = 17n + 54 words (tendency 2n, RK update +
    divergence 4n+1 (3n+1 on stage 1), 5-pass FFT Poisson 10, correction 7 per substage)."""
    return (17 * n_prognostic + 54) * word


def run_gpu(args):
    import oceananigans_mi355x as om
    from oceananigans_mi355x import workloads

    N = tuple(int(x) for x in args.size.split(","))
    ft = np.float32 if args.precision == "fp32" else np.float64
    word = 4 if ft == np.float32 else 8
    grid, kw, init = workloads.free_convection(N=N, float_type=ft)
    model = om.NonhydrostaticModel(grid, **kw)
    om.set_(model, **init)
    dt = 1.0
    if args.warmup:
        model.time_step(dt, args.warmup)
    if args.kchunk:                          # fixed (profiling runs: no autotuner launches in the trace)
        model.set_tuning(kchunk=args.kchunk)
        kchunk = args.kchunk
    else:
        kchunk = model.autotune()            # z-chunk length of the tiled kernels, timed on this state
    # per-phase device times: HIP events around every phase (this run issues every launch directly)
    model.set_phase_timing(True)
    model.time_step(dt, args.steps)
    phases = model.phase_times()
    phased_ms = model.last_step_ms()
    model.set_phase_timing(False)
    # the headline: time_step as a caller runs it, pairs of steps replaying the model's captured
    # two-step hipGraph (captured by the untimed pair first)
    model.set_tuning(step_graphs=not args.no_step_graphs)
    model.time_step(dt, 2)
    t0 = time.perf_counter()
    model.time_step(dt, args.steps)          # synchronous: returns after the device has finished
    t1 = time.perf_counter()
    ev_ms = model.last_step_ms()
    wall = t1 - t0
    ncell = N[0] * N[1] * N[2]
    n = 3 + N_TRACERS
    bytes_step = algorithmic_bytes_per_cell_step(n, word) * ncell
    # dominant kernel: the phase with the largest device time
    dom = max(phases, key=lambda p: phases[p][0])
    dom_ms, dom_launches = phases[dom]
    per_cell = {"tendencies": 2 * n * word,                       # read n fields, write n tendencies
                "poisson": 10 * word, "correction": 7 * word,
                "rk_update": (4 * n + 1) * word, "closure": (3 + N_TRACERS + 1 + N_TRACERS) * word}.get(dom, 2 * n * word)
    dom_bytes = per_cell * ncell
    dom_avg_s = dom_ms / max(dom_launches, 1) / 1e3
    achieved = dom_bytes / dom_avg_s / 1e9 if dom_avg_s > 0 else 0.0
    div = float(np.abs(model.divergence()).max())
    finite = all(np.isfinite(model.velocities[v].interior()).all() for v in "uvw")
    model.close()
    return dict(N=N, kchunk=kchunk, ncell=ncell, wall=wall, ev_ms=ev_ms, phased_ms=phased_ms, phases=phases, dom=dom,
                dom_avg_ms=dom_avg_s * 1e3,
                dom_bytes=dom_bytes, achieved=achieved, bytes_step=bytes_step, word=word, div=div, finite=finite)


def host_cpu():
    """(CPU model, physical cores usable by this process) from lscpu and the CPU affinity mask."""
    import subprocess
    model, cores = "unknown", None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
        allowed = os.sched_getaffinity(0)
        table = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True, timeout=20).stdout
        phys = {(int(c), int(s)) for cpu, c, s in (l.split(",") for l in table.splitlines() if l and l[0] != "#")
                if int(cpu) in allowed}
        cores = len(phys)
    except Exception:
        pass
    return model, cores or len(os.sched_getaffinity(0))


def run_cpu_baseline(args):
    """The FP64 oracle (a CPU restatement of the reference algorithm, test infrastructure, not the
    product) timed on this host per benchmarking/src/utils.jl:84-101: one warm-up step, then timed
    steps.  All physical cores usable by the process on the benchmark's full size; one thread on a
    bounded sample of the same workload (the per-cell cost of a step is size-independent to ~10%)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from cases import oracle_for
    from oracle import oracle as O
    from oceananigans_mi355x import workloads

    model_name, phys = host_cpu()
    cap = int(os.environ.get("OMP_NUM_THREADS", phys) or phys)
    threads = max(1, min(phys, cap))

    def timed(N, nthreads, steps):
        O.set_threads(nthreads)
        grid, kw, init = workloads.free_convection(N=N, float_type=np.float64)
        orc = oracle_for(grid, kw, init)
        orc.time_step(1.0, 1)                     # warm-up step
        t0 = time.perf_counter()
        orc.time_step(1.0, steps)
        dt = time.perf_counter() - t0
        orc.close()
        return N[0] * N[1] * N[2] * steps / dt, dt

    Nfull = tuple(int(x) for x in args.size.split(","))
    full, t_full = timed(Nfull, threads, args.cpu_steps)
    Ns = tuple(int(x) for x in args.cpu_sample.split(","))
    one, t_one = timed(Ns, 1, 2)
    return {"value": full, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle (CPU restatement of the reference algorithm), the benchmark's physics at "
                      f"{'x'.join(map(str, Nfull))}, {args.cpu_steps} RK3 step(s) after 1 warm-up step in {t_full:.1f} s, "
                      f"OpenMP threads = {threads} (the physical cores of this process's CPU set)",
            "cpu_model": model_name, "physical_cores": phys,
            "one_thread": {"value": one, "unit": "cell-updates/s",
                           "sample": f"{'x'.join(map(str, Ns))}, 2 RK3 steps after 1 warm-up step in {t_one:.1f} s, 1 thread"}}


def run_cg_seamount(args):
    """Pressure solves on the immersed, stretched seamount (FP32 model, FP64 CG iterate and
    residual, reltol 1e-8): FFT-preconditioned CG against flexible CG with the learned correction.
    The network is trained on the GPU first (rounds of one logged time step and Adam steps); then
    both models start from the same state and the device time of one time step's three solves is
    compared, with the iteration counts and the held-out A-energy gain."""
    import oceananigans_mi355x as om
    from oceananigans_mi355x import workloads

    N = tuple(int(x) for x in args.cg_size.split(","))
    g_fft, kw_fft, init = workloads.seamount(N=N, preconditioner=om.FourierTridiagonalPoissonSolver())
    g_ln, kw_ln, _ = workloads.seamount(N=N, preconditioner=om.LearnedPreconditioner())
    fft = om.NonhydrostaticModel(g_fft, **kw_fft)
    om.set_(fft, **init)
    learned = om.NonhydrostaticModel(g_ln, **kw_ln)
    om.set_(learned, **init)
    rng = np.random.default_rng(7)
    w = (rng.normal(0, np.sqrt(2.0 / 432), size=(7, 27, 16, 16))).astype(np.float32)
    w[0, :, 3:, :] = 0
    w[6] = 0                                     # zero output layer: starts as the FFT preconditioner
    learned.learned_set_weights(w)
    # online training over the first cg_rounds time steps: each step's solves log (input, target)
    # pairs from their first cg_log_iterations iterations into a ring of cg_pairs, then Adam steps
    # run over the ring
    learned.learned_log(iterations_per_solve=args.cg_log_iterations, capacity=args.cg_pairs)
    t0 = time.perf_counter()
    for rnd in range(args.cg_rounds):
        learned.time_step(args.cg_dt, 1)
        l0, l1 = learned.learned_train(args.cg_train_steps, lr=args.cg_lr)
        if rnd % 20 == 0 or rnd == args.cg_rounds - 1:
            print(f"bench.py: seamount training step {rnd}: energy loss {l0:.3f} -> {l1:.3f}, "
                  f"{learned.learned_pairs_logged()} pairs logged", file=sys.stderr, flush=True)
    t_train = time.perf_counter() - t0
    pairs_logged = learned.learned_pairs_logged()
    learned.learned_log(iterations_per_solve=8, capacity=24)
    if args.cg_learned_iterations:
        learned.set_tuning(learned_iterations=args.cg_learned_iterations)
    state = om.prognostic_state(learned)
    om.restore_prognostic_state_(fft, state)
    res = {}
    for name, m in (("fft", fft), ("learned", learned)):
        m.set_phase_timing(True)
        m.time_step(args.cg_dt, 1)
        ph = m.phase_times()
        it, r, true_r = m.last_solve_info()
        res[name] = {"ms_per_solve": ph["poisson"][0] / 3, "iterations_last_solve": it, "recomputed_residual": true_r,
                     "step_phases_ms": {k: round(v[0], 4) for k, v in ph.items()}}
    res["learned"]["held_out_energy_ratio"] = learned.learned_evaluate()
    res["learned"]["network_iterations_last_solve"], res["learned"]["fell_back"] = learned.learned_last_solve()
    fft.close()
    learned.close()
    return {"grid": "x".join(map(str, N)), "reltol": 1e-8, "fft": res["fft"], "learned": res["learned"],
            "learned_over_fft_time": res["learned"]["ms_per_solve"] / res["fft"]["ms_per_solve"],
            "dt": args.cg_dt, "training": {"time_steps": args.cg_rounds, "adam_steps_per_time_step": args.cg_train_steps, "lr": args.cg_lr,
                         "pairs_logged": pairs_logged, "pair_ring_capacity": args.cg_pairs,
                         "iterations_logged_per_solve": args.cg_log_iterations, "wall_s": t_train},
            "network_iterations_per_solve": args.cg_learned_iterations or "all"}


def run_ensemble(args):
    """Batched ensemble of uneven-size free-convection LES (FP32, WENO5 + AMD), after the autotuner
    has picked each size's z-chunk length: all members advanced together in the batched form (one
    launch per kernel of the step over the table of all members, each pair of steps one hipGraph
    replay) for the configuration's 500 steps, against members advanced one after another
    (om_time_step per member, as the reference would) and against per-member graphs on concurrent
    streams, each timed on `--ensemble-compare-steps` steps.  Total cell-updates/s of each."""
    import oceananigans_mi355x as om
    from oceananigans_mi355x import workloads

    t0 = time.perf_counter()
    cfgs = workloads.ensemble_members(n=args.ensemble_members, seed=5)
    members = []
    for grid, kw, init in cfgs:
        m = om.NonhydrostaticModel(grid, **kw)
        om.set_(m, **init)
        members.append(m)
    print(f"bench.py: ensemble of {len(members)} members created in {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    tuned = {}
    for m in members:                       # one autotune per distinct size
        key = m.grid.N
        if key not in tuned:
            m.time_step(1.0, 1)
            tuned[key] = m.autotune()
        else:
            m.set_tuning(kchunk=tuned[key])
    ncell = sum(int(np.prod(m.grid.N)) for m in members)
    steps, cmp_steps, dt = args.ensemble_steps, args.ensemble_compare_steps, 1.0
    for m in members:                       # warm-up: first steps and captured graphs
        m.time_step(dt, 1)
    ens = om.Ensemble(members)
    ens.time_step(dt, 4, mode="member_graphs")
    ens.time_step(dt, 4, mode="batched")
    t0 = time.perf_counter()
    for m in members:
        m.time_step(dt, cmp_steps)
    t_seq = time.perf_counter() - t0
    t0 = time.perf_counter()
    ens.time_step(dt, cmp_steps, mode="member_graphs")
    t_graphs = time.perf_counter() - t0
    by_groups = {}
    for groups in (1, 2, 4, 8):             # member groups of the batched form, each graph on its own stream
        ens.set_batch_groups(groups)
        ens.time_step(dt, 2, mode="batched")  # records and captures
        t0 = time.perf_counter()
        ens.time_step(dt, cmp_steps, mode="batched")
        by_groups[groups] = ncell * cmp_steps / (time.perf_counter() - t0)
    best = max(by_groups, key=by_groups.get)
    print(f"bench.py: ensemble compared over {cmp_steps} steps; batched run of {steps} steps in {best} groups next",
          file=sys.stderr, flush=True)
    ens.set_batch_groups(best)
    ens.time_step(dt, 2, mode="batched")
    t0 = time.perf_counter()
    ens.time_step(dt, steps, mode="batched")
    t_ens = time.perf_counter() - t0
    finite = all(not m.velocities[n].hasnan() for m in members for n in ("u", "v", "w"))
    sizes = {}
    for m in members:
        k = "x".join(map(str, m.grid.N))
        sizes[k] = sizes.get(k, 0) + 1
    ens.close()
    for m in members:
        m.close()
    return {"members": len(members), "sizes": sizes, "cells": ncell, "steps": steps,
            "kchunk_by_size": {"x".join(map(str, k)): v for k, v in tuned.items()},
            "batched_cell_updates_per_s": ncell * steps / t_ens, "batched_wall_s": t_ens, "batched_groups": best,
            "batched_cell_updates_per_s_by_groups": by_groups,
            "compare_steps": cmp_steps,
            "one_after_another_cell_updates_per_s": ncell * cmp_steps / t_seq,
            "member_graphs_cell_updates_per_s": ncell * cmp_steps / t_graphs,
            "batched_over_one_after_another": (ncell * steps / t_ens) / (ncell * cmp_steps / t_seq),
            "finite": bool(finite)}


def run_mixed(args):
    """The mixed-precision wind-and-convection LES at full size (512×512×256, WENO5, SmagorinskyLilly,
    linear EOS with T and S, surface momentum and heat fluxes; tendencies and FFTs in FP32): device
    time per step without and with the FP64 iterative-refinement residual on every pressure solve."""
    import oceananigans_mi355x as om
    from oceananigans_mi355x import workloads

    N = tuple(int(x) for x in args.mixed_size.split(","))
    grid, kw, init = workloads.wind_convection(N=N, float_type=np.float32)
    model = om.NonhydrostaticModel(grid, **kw)
    om.set_(model, **init)
    model.time_step(1.0, 2)
    kchunk = model.autotune()
    res = {"grid": "x".join(map(str, N)), "kchunk": kchunk, "steps": args.mixed_steps}
    for name, refine in (("fp32", False), ("fp32_fp64_refinement", True)):
        model.set_tuning(poisson_refinement=refine)
        model.time_step(1.0, 1)
        model.set_phase_timing(True)
        t0 = time.perf_counter()
        model.time_step(1.0, args.mixed_steps)
        wall = time.perf_counter() - t0
        ph = model.phase_times()
        model.set_phase_timing(False)
        res[name] = {"ms_per_step": wall * 1e3 / args.mixed_steps,
                     "cell_updates_per_s": N[0] * N[1] * N[2] * args.mixed_steps / wall,
                     "poisson_ms_per_step": ph["poisson"][0] / args.mixed_steps}
    w = model.velocities.w.interior()
    res["finite"] = bool(np.isfinite(w).all())
    res["max_abs_divergence"] = float(np.abs(model.divergence()).max())
    res["max_abs_w"] = float(np.abs(w).max())
    model.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", default="256,256,128")
    ap.add_argument("--precision", choices=("fp32", "fp64"), default="fp32")
    ap.add_argument("--cpu-sample", default="64,64,64", help="the one-thread CPU baseline's sample size")
    ap.add_argument("--cpu-steps", type=int, default=1, help="timed steps of the all-cores CPU baseline (full size)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cg", action="store_true", help="skip the seamount CG measurement (learned vs FFT preconditioner)")
    ap.add_argument("--cg-size", default="256,256,128")
    ap.add_argument("--cg-rounds", type=int, default=200, help="time steps of online training (one logged step each)")
    ap.add_argument("--cg-train-steps", type=int, default=10, help="Adam steps after each logged time step")
    ap.add_argument("--cg-pairs", type=int, default=96, help="capacity of the ring of logged pairs")
    ap.add_argument("--cg-log-iterations", type=int, default=3, help="leading iterations of each solve that log a pair")
    ap.add_argument("--cg-lr", type=float, default=1e-4, help="Adam learning rate of the online training")
    ap.add_argument("--cg-dt", type=float, default=10.0,
                    help="time step of the seamount run (s); its flow speeds up near the seamount, to ~1 m/s by t = 2000 s")
    ap.add_argument("--cg-learned-iterations", type=int, default=2,
                    help="leading CG iterations that apply the network (0: all)")
    ap.add_argument("--kchunk", type=int, default=0, help="z-chunk length of the tiled kernels (0: autotune)")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the uneven-size ensemble measurement")
    ap.add_argument("--no-step-graphs", action="store_true",
                    help="time the headline steps with every launch issued directly (no captured graph)")
    ap.add_argument("--no-mixed", action="store_true", help="skip the 512x512x256 mixed-precision wind-and-convection run")
    ap.add_argument("--mixed-size", default="512,512,256")
    ap.add_argument("--mixed-steps", type=int, default=10)
    ap.add_argument("--ensemble-members", type=int, default=64)
    ap.add_argument("--ensemble-steps", type=int, default=500)
    ap.add_argument("--ensemble-compare-steps", type=int, default=40)
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="only run a plain device copy of this many MiB (counter calibration) and exit")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "traffic_free_convection_256x256x128_fp32.json"),
                    help="measured HBM bytes per launch of the dominant kernel (from the FETCH_SIZE/WRITE_SIZE passes)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": 4 * n}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is design-only)", file=sys.stderr)
        sys.exit(2)

    r = run_gpu(args)
    ms_per_step = r["wall"] * 1e3 / args.steps
    value = r["ncell"] * args.steps / r["wall"]
    cpu = None if args.no_cpu_baseline else run_cpu_baseline(args)
    cg = None
    if not args.no_cg:
        try:
            cg = run_cg_seamount(args)
        except Exception as e:                       # reported, not hidden: the main metric stands alone
            cg = {"error": f"{type(e).__name__}: {e}"}
    ens = None
    if not args.no_ensemble:
        try:
            ens = run_ensemble(args)
        except Exception as e:                       # reported, not hidden
            ens = {"error": f"{type(e).__name__}: {e}"}
    mixed = None
    if not args.no_mixed:
        try:
            mixed = run_mixed(args)
        except Exception as e:                       # reported, not hidden
            mixed = {"error": f"{type(e).__name__}: {e}"}
    step_gbs = r["bytes_step"] * args.steps / r["wall"] / 1e9
    traffic = None
    valu_issue = None
    if os.path.exists(args.traffic):
        t = json.load(open(args.traffic))
        if (t.get("kernel") == r["dom"] and t.get("grid") == "x".join(map(str, r["N"]))
                and t.get("precision") == args.precision):
            traffic = t["hbm_bytes_per_launch"]
            if t.get("sq_insts_valu_per_launch") and r["dom_avg_ms"] > 0:
                # the dominant kernel is issue-bound: its measured VALU wave-instructions per launch
                # over what 256 CUs issue in the launch's duration at one per CU per cycle (2.4 GHz)
                valu_issue = t["sq_insts_valu_per_launch"] / (256 * 2.4e9 * r["dom_avg_ms"] * 1e-3)
    out = {
        "metric": "cell-updates/s per RK3 step (256³-class FP32 LES: 256x256x128 WENO5+AMD free convection)",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if r["word"] == 4 else "f64",
        "data": "synthetic (seeded SplitMix64 initial buoyancy noise; the reference's configuration)",
        "config": {"workload": "free_convection_LES", "grid": "x".join(map(str, r["N"])),
                   "advection": "WENO5", "closure": "AnisotropicMinimumDissipation", "buoyancy": "BuoyancyTracer",
                   "coriolis": "FPlane(f=1e-4)", "pressure_solver": "FFT", "dt": 1.0, "parallelism": "single GPU",
                   "kchunk": r["kchunk"]},
        "roofline": {"bound": "hbm", "achieved": r["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": r["achieved"] / HBM_PEAK_GBS, "traffic": traffic, "kernel": r["dom"],
                     "kernel_avg_ms": r["dom_avg_ms"], "algorithmic_bytes_per_launch": r["dom_bytes"],
                     "valu_issue_frac": valu_issue},
        "step_roofline": {"algorithmic_bytes_per_cell_step": algorithmic_bytes_per_cell_step(3 + N_TRACERS, r["word"]),
                          "achieved_GBs": step_gbs, "frac": step_gbs / HBM_PEAK_GBS},
        "phases_ms_per_step": {k: v[0] / args.steps for k, v in r["phases"].items()},
        "device_ms_per_step": r["ev_ms"] / args.steps,
        "step_graphs": not args.no_step_graphs,
        "phase_timed_run_ms_per_step": r["phased_ms"] / args.steps,
        "max_abs_divergence": r["div"],
        "finite": r["finite"],
        "cpu_baseline": cpu,
        "cg_seamount": cg,
        "ensemble": ens,
        "mixed_precision_wind_convection": mixed,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
