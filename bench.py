"""Benchmark: RK4 time stepping of the 513x513x9 Gaussian-shoal tank (6th order, 5-wavelength
sponges, 64-component JONSWAP initial surface), mixed precision (FP32 multigrid preconditioner,
FP64 outer solve / matvec / RK4 state) against the plain FP64 GPU form, on one MI355X.

Prints one JSON line.  value = DOF-steps/s of the mixed-precision run (DOF = Nx*Ny*(Nz+1), the
reference's count, OceanWave3D.f90:29).  The timed region covers exactly --steps RK4 steps with the
state resident in HBM, bracketed by device synchronisation.
"""
import argparse
import collections
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "oceanwave3d-fortran90_amd"))

WORKLOAD = "shoal513"
HBM_PEAK_GBS = 8000.0


MG_LEVELS = None   # None: the autotuner's cached schedule for the class, else 7 levels


def tank_cfg(cfg, mg_precision, solver=1, gmres_side=0):
    """The 513^2 tank's solver settings and, unless --mg-levels is given, the autotuner's cached schedule
    (ow3d/tuning.json) for its class; without a cached entry 7 levels (513^2 .. 9^2 x 9, whose 1,210
    unknowns are solved exactly) and the default launch choices."""
    from ow3d import autotune
    cfg = dict(cfg, mg_precision=mg_precision, solver=solver, reltol=1e-8, abstol=0.0, gmres_m=23, gmres_side=gmres_side)
    # FP16 coarse storage (mg_precision 2) uses the FP32 multigrid's schedule
    choice = autotune.tuned(dict(cfg, mg_precision=min(mg_precision, 1))) if MG_LEVELS is None else None
    cfg["mg_maxgrids"] = choice["mg_maxgrids"] if choice else (MG_LEVELS or 7)
    return cfg, choice


def make_tank(ow3d, problems, mg_precision, solver=1, seed=0, gmres_side=0):
    from ow3d import autotune
    cfg, bath, E, P, dt, zones = problems.workload(WORKLOAD, seed=seed)
    cfg, choice = tank_cfg(cfg, mg_precision, solver, gmres_side)
    T = ow3d.Tank(cfg, **bath)
    autotune.apply(T, choice)
    for z in zones:
        T.add_relax_zone(*z)
    T.set_state(E, P)
    return T, cfg, dt


def run_steps(T, steps, dt, t0=0.0, status=None):
    """Per-step wall times of `steps` RK4 steps; status (a list) receives each step's return code
    (1: a Laplace solve of that step missed its tolerance)."""
    lat = []
    for s in range(steps):
        a = time.perf_counter()
        st = T.rk4_steps(1, t0 + s * dt, dt)
        lat.append(time.perf_counter() - a)
        if status is not None:
            status.append(st)
    return lat


def _oracle_tank(workload):
    """The workload's tank on the oracle in reference-faithful mode: CSR level matrices with the
    cross-derivative replication, lexicographic Gauss-Seidel with the special-point passes, unfused
    derivative passes, MGS GMRES(23) at rtol 1e-8."""
    sys.path.insert(0, ROOT)
    import oracle
    from ow3d import problems
    cfg, bath, E, P, dt, zones = problems.workload(workload)
    p = dict(nx=cfg["nx"], ny=cfg["ny"], nz=cfg["nz"], alpha=3, beta=3, gamma=3, gridz=1, lx=cfg["lx"], ly=cfg["ly"],
             lz=1.0, g=9.81, nonlinear=1, solver=1, precond=3, gmres_m=23, reltol=1e-8, abstol=0.0, maxit=1, nu1=1,
             nu2=1, mg_maxgrids=20, extrapolation=1, mg_mode=0, cd_quirk=1)
    T = oracle.OracleTank(p, bath["h"], bath["hx"], bath["hy"], bath["hxx"], bath["hyy"])
    for z in zones:
        T.add_relax_zone(*z)
    T.set_state(E, P)
    return T, cfg, dt


def _cpu_steps(workload, steps):
    """Wall seconds and DOF of `steps` RK4 steps of the workload's tank on the oracle (one host thread);
    run in a child process for the one-tank-per-core figure."""
    T, cfg, dt = _oracle_tank(workload)
    t0 = time.perf_counter()
    for s in range(steps):
        T.rk4_step(s * dt, dt)
    el = time.perf_counter() - t0
    T.close()
    return el, cfg["nx"] * cfg["ny"] * (cfg["nz"] + 1) * steps


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(per_core=True, max_procs=16):
    """The oracle in reference-faithful mode, one host thread per tank: one RK4 step of the 513x513x9
    shoal tank itself (value), then, with per_core, one tank per host core (at most max_procs processes,
    the box's CPU share) each taking one RK4 step of the 257x257x9 member of the same family."""
    import multiprocessing as mp
    el, dof_steps = _cpu_steps(WORKLOAD, 1)
    out = dict(value=dof_steps / el, unit="DOF-steps/s", cores=1, kind="port",
               sample=f"1 RK4 step of the 513x513x9 Gaussian-shoal tank (the benchmark workload), reference-faithful "
                      f"oracle (CSR multigrid, lexicographic Gauss-Seidel, GMRES(23)), {el:.1f} s on 1 host core",
               step_s=el, cpu_model=_cpu_model(), nproc=os.cpu_count())
    if per_core:
        procs = max(1, min(max_procs, os.cpu_count() or 1))
        ctx = mp.get_context("spawn")   # the children never touch the GPU
        t0 = time.perf_counter()
        with ctx.Pool(procs) as pool:
            res = pool.starmap(_cpu_steps, [("shoal257", 1)] * procs)
        wall = time.perf_counter() - t0
        slowest = max(r[0] for r in res)
        out["one_tank_per_core"] = dict(
            processes=procs, dof_steps_per_s=sum(r[1] for r in res) / slowest, slowest_step_s=slowest,
            wall_s_incl_setup=wall,
            sample=f"{procs} processes, each 1 RK4 step of its own 257x257x9 Gaussian-shoal tank (same family and "
                   f"resolution as the 513 workload); throughput = all tanks' DOF-steps / slowest step")
    return out


def main():
    global MG_LEVELS
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--err-steps", type=int, default=200,
                    help="RK4 steps of the mixed-vs-FP64 error leg on the 513 workload (0: skip)")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 HBM-counter runs of the matvec")
    ap.add_argument("--profile-precond", type=int, default=0, metavar="K",
                    help="only apply the FP32 multigrid preconditioner K times to a random vector, then exit")
    ap.add_argument("--mg-levels", type=int, default=None,
                    help="multigrid levels of the 513^2 tank (MGmaxgrids); default: the autotuner's cached choice, else 7")
    ap.add_argument("--batched-tanks", type=int, default=2,
                    help="also advance this many 513x513x9 tanks together (different JONSWAP seeds); 0: skip")
    ap.add_argument("--sweep-threads", type=int, default=None,
                    help="with --profile-precond: zebra sweep launch choice instead of the autotuner's (see ow3d_tune)")
    ap.add_argument("--batched-only", action="store_true",
                    help="only the --batched-tanks run of the 513x513x9 tanks (prints its own JSON line)")
    ap.add_argument("--batch", type=int, default=0, metavar="N",
                    help="many-tank workload instead: N uneven tanks advanced together (prints its own JSON line)")
    ap.add_argument("--untuned", action="store_true", help="--batch / --serve without the autotuner's cached schedules")
    ap.add_argument("--serve", type=float, default=0.0, metavar="SECONDS",
                    help="request-loop workload instead: Poisson arrivals served for SECONDS (prints its own JSON line)")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="with --serve: Poisson arrival rate per second")
    ap.add_argument("--profile-steps", type=int, default=0, metavar="K",
                    help="only run 2 + K RK4 steps of the 513^2 workload with the multigrid precision of "
                         "--precision, then exit (for rocprofv3 kernel traces)")
    ap.add_argument("--precision", type=int, default=1, help="with --profile-steps: 0 FP64, 1 FP32, 2 FP16 coarse")
    ap.add_argument("--profile-matvec", type=int, default=0, metavar="K",
                    help="only run K high-order matvec launches of the workload and a 256 MiB device copy "
                         "(the calibration kernel for HBM counters), then exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported"}))
        sys.exit(2)
    # HBM-counter runs first, as child processes, before this process touches the GPU; then the CPU baseline
    # (host only, its one-tank-per-core processes started before the GPU is initialised)
    side = args.profile_matvec or args.profile_precond or args.profile_steps or args.batch or args.serve or args.batched_only
    tr = None if (args.no_traffic or side) else measure_traffic()
    cpu = None if (args.no_cpu or side) else cpu_baseline()
    import ow3d
    from ow3d import problems
    MG_LEVELS = args.mg_levels
    if args.profile_matvec:
        profile_matvec(ow3d, problems, args.profile_matvec)
        return
    if args.batch:
        print(json.dumps(bench_batch(ow3d, problems, args.batch, args.steps, args.warmup, use_tuning=not args.untuned)),
              flush=True)
        return
    if args.serve:
        print(json.dumps(bench_serve(args.serve, rate=args.serve_rate, tuned=not args.untuned)), flush=True)
        return
    if args.batched_only:
        print(json.dumps(bench_batched_513(ow3d, problems, args.batched_tanks, args.steps, args.warmup)), flush=True)
        return
    if args.profile_steps:
        T, cfg, dt = make_tank(ow3d, problems, args.precision)
        run_steps(T, 2, dt)
        st = []
        lat = run_steps(T, args.profile_steps, dt, 2 * dt, st)
        print(json.dumps({"profile_steps": args.profile_steps, "precision": args.precision,
                          "ms_per_step": 1e3 * float(np.mean(lat)), "status": st}), flush=True)
        T.close()
        return
    if args.profile_precond:
        T, cfg, dt = make_tank(ow3d, problems, 1)
        if args.sweep_threads is not None:
            from ow3d import autotune
            ch = autotune.tuned(cfg)
            T.tune(ch["matvec_variant"] if ch else 0, args.sweep_threads)
        r = np.random.default_rng(0).standard_normal(T.shape3)
        for _ in range(args.profile_precond):
            T.precond_apply(r)
        print(json.dumps({"precond_applications": args.profile_precond}), flush=True)
        T.close()
        return

    results = {}
    for label, prec, side in (("fp64", 0, 0), ("mixed", 1, 0), ("mixed_fp16_coarse", 2, 0), ("mixed_flexible", 1, 1)):
        T, cfg, dt = make_tank(ow3d, problems, prec, gmres_side=side)
        run_steps(T, args.warmup, dt)
        T.set_state(*make_tank_state(problems))
        status = []
        lat = run_steps(T, args.steps, dt, status=status)
        E, Pp, W, phi = T.get_state()
        tot, solves, itmin, itmax = T.stats()
        replays, captures, capture_ms = T.step_stats()
        results[label] = dict(lat=lat, E=E, P=Pp, iters=(tot, solves, itmin, itmax), unconverged_steps=sum(status),
                              graphs=dict(replays=replays, captures=captures, capture_ms=round(capture_ms, 2)))
        if label == "mixed":
            T.time_matvec(5)
            mv_ms = T.time_matvec(200)
            ncol = T.nxg * T.nyg
            b_mv = 16 * T.n + 8 * 5 * ncol
            from ow3d import autotune
            ch = autotune.tuned(cfg) if MG_LEVELS is None else None
            mv_kernel = {2: "k_matvec3p", 3: "k_matvec3s"}.get(ch["matvec_variant"] if ch else 0, "k_matvec3c")
        T.close()
    dof = cfg["nx"] * cfg["ny"] * (cfg["nz"] + 1)
    mix, f64 = results["mixed"], results["fp64"]
    batched = bench_batched_513(ow3d, problems, args.batched_tanks, args.steps, args.warmup) if args.batched_tanks > 1 else None
    t_mix = sum(mix["lat"])
    t_f64 = sum(f64["lat"])
    err_eta = float(np.max(np.abs(mix["E"] - f64["E"])) / np.max(np.abs(f64["E"])))
    err_phi = float(np.max(np.abs(mix["P"] - f64["P"])) / np.max(np.abs(f64["P"])))
    achieved = b_mv / (mv_ms * 1e-3) / 1e9
    out = {
        "metric": "DOF-steps/s mixed vs FP64 (<=1e-6 rel eta err); p50/p99 RK4-step latency",
        "value": dof * args.steps / t_mix,
        "unit": "DOF-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_mix / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (FP32 multigrid preconditioner)",
        "data": "synthetic",
        "config": {"workload": "513x513x9 Gaussian shoal, 6th order, 5-wavelength sponges, JONSWAP seed 0",
                   "dof": dof, "solver": f"GMRES(23) rtol 1e-8 + V(1,1) zebra-line multigrid, {cfg['mg_maxgrids']} levels (mixed run)"},
        "fp64": {"value": dof * args.steps / t_f64, "ms_per_step": 1e3 * t_f64 / args.steps,
                 "p50_ms": 1e3 * float(np.percentile(f64["lat"], 50)), "p99_ms": 1e3 * float(np.percentile(f64["lat"], 99)),
                 "iters": f64["iters"], "step_graphs": f64["graphs"], "unconverged_steps": f64["unconverged_steps"]},
        "mixed": {"p50_ms": 1e3 * float(np.percentile(mix["lat"], 50)), "p99_ms": 1e3 * float(np.percentile(mix["lat"], 99)),
                  "iters": mix["iters"], "step_graphs": mix["graphs"], "unconverged_steps": mix["unconverged_steps"],
                  "rel_err_eta_vs_fp64": err_eta,
                  "rel_err_phi_vs_fp64": err_phi},
        "speedup_mixed_vs_fp64": t_f64 / t_mix,
        "roofline": {"bound": "hbm", "kernel": f"{mv_kernel} (high-order sigma-Laplace matvec)", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "avg_launch_us": mv_ms * 1e3, "algorithmic_bytes": b_mv},
    }
    f16 = results["mixed_fp16_coarse"]
    out["mixed_fp16_coarse"] = {
        "value": dof * args.steps / sum(f16["lat"]), "ms_per_step": 1e3 * sum(f16["lat"]) / args.steps,
        "p50_ms": 1e3 * float(np.percentile(f16["lat"], 50)), "iters": f16["iters"],
        "rel_err_eta_vs_fp64": float(np.max(np.abs(f16["E"] - f64["E"])) / np.max(np.abs(f64["E"]))),
        "note": "FP32 multigrid with the coarsest level's dense inverse stored in FP16 (FP32 sums)"}
    fl = results["mixed_flexible"]
    out["mixed_flexible"] = {
        "value": dof * args.steps / sum(fl["lat"]), "ms_per_step": 1e3 * sum(fl["lat"]) / args.steps,
        "p50_ms": 1e3 * float(np.percentile(fl["lat"], 50)), "iters": fl["iters"], "unconverged_steps": fl["unconverged_steps"],
        "rel_err_eta_vs_fp64": float(np.max(np.abs(fl["E"] - f64["E"])) / np.max(np.abs(f64["E"]))),
        "rel_err_phi_vs_fp64": float(np.max(np.abs(fl["P"] - f64["P"])) / np.max(np.abs(f64["P"]))),
        "note": "FP32 multigrid with flexible right-preconditioned GMRES (stops on the FP64 true residual)"}
    if batched is not None:
        batched["mixed_batched_vs_plain_fp64"] = batched["mixed"]["dof_steps_per_s"] / out["fp64"]["value"]
        out["batched"] = batched
    if tr is not None:
        out["roofline"]["traffic"] = tr["bytes"]
        out["roofline"]["traffic_detail"] = tr
    if args.err_steps > 0:
        out["err_leg"] = error_leg(ow3d, problems, args.err_steps)
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


def error_leg(ow3d, problems, steps):
    """The 513 workload stepped `steps` times (the configuration's 200) with the FP64 multigrid and with the
    FP32 multigrid from the same initial state: relative L-inf difference of eta and Phi at the end, steps
    in which a solve missed its tolerance, and the wall time of each run (one call per step)."""
    res = {}
    for label, prec in (("fp64", 0), ("mixed", 1)):
        T, cfg, dt = make_tank(ow3d, problems, prec)
        status = []
        lat = run_steps(T, steps, dt, status=status)
        E, P, W, phi = T.get_state()
        res[label] = dict(E=E, P=P, unconverged_steps=sum(status), wall_s=sum(lat), iters=T.stats())
        T.close()
    f, m = res["fp64"], res["mixed"]
    return {"steps": steps,
            "rel_err_eta_vs_fp64": float(np.max(np.abs(m["E"] - f["E"])) / np.max(np.abs(f["E"]))),
            "rel_err_phi_vs_fp64": float(np.max(np.abs(m["P"] - f["P"])) / np.max(np.abs(f["P"]))),
            "unconverged_steps": {k: v["unconverged_steps"] for k, v in res.items()},
            "wall_s": {k: round(v["wall_s"], 3) for k, v in res.items()},
            "iters": {k: v["iters"] for k, v in res.items()}}


def bench_batched_513(ow3d, problems, k, steps, warmup):
    """k tanks of the 513x513x9 workload (JONSWAP seeds 0 .. k-1) advanced together with
    ow3d.batch_rk4_steps, FP64 and FP32 multigrid: the latency of a step of the whole batch (one call per
    step, so every step waits for the slowest tank), then DOF-steps/s of the whole batch over `steps` more
    steps in one call (each tank starts its next step as soon as its previous one has settled; the stage
    graphs the step counts need are captured by then)."""
    out = {"tanks": k}
    for label, prec in (("fp64", 0), ("mixed", 1)):
        tanks, dts = [], []
        for s in range(k):
            T, cfg, dt = make_tank(ow3d, problems, prec, seed=s)
            tanks.append(T); dts.append(dt)
        ow3d.batch_rk4_steps(tanks, warmup, dts)
        lat = []
        for _ in range(steps):
            a = time.perf_counter()
            ow3d.batch_rk4_steps(tanks, 1, dts)
            lat.append(time.perf_counter() - a)
        a = time.perf_counter()
        ow3d.batch_rk4_steps(tanks, steps, dts)
        t_all = time.perf_counter() - a
        dof = cfg["nx"] * cfg["ny"] * (cfg["nz"] + 1) * k
        out[label] = {"dof_steps_per_s": dof * steps / t_all, "ms_per_batch_step": 1e3 * t_all / steps,
                      "dof_steps_per_s_stepwise": dof * steps / sum(lat),
                      "p50_ms": 1e3 * float(np.percentile(lat, 50)), "p99_ms": 1e3 * float(np.percentile(lat, 99))}
        for T in tanks:
            T.close()
    return out


def bench_batch(ow3d, problems, ntanks, steps, warmup, seed=0, use_tuning=True):
    """Many-tank workload: ntanks tanks drawn as problems.batch_draw (Nx, Ny in {65..513}, Nz in {6, 9, 12},
    order 4/6/8, own depth, spectrum and seed), advanced together (ow3d.batch_rk4_steps: each tank on its
    own stream, steps replayed from captured graphs) and then one tank at a time, FP64 and FP32 multigrid.
    Reports DOF-steps/s of both and the largest eta difference between the precisions."""
    from ow3d import autotune
    table = autotune.load() if use_tuning else {}
    draws = [problems.batch_draw(seed * 1000003 + i) for i in range(ntanks)]
    res = {}
    n_tuned = 0
    for label, prec in (("fp64", 0), ("mixed", 1)):
        tanks, dts, dof = [], [], 0
        t_build = time.perf_counter()
        for d in draws:
            cfg, bath, E, P, dt, _ = problems.batch_tank_from(d)
            # FP32 multigrid with flexible GMRES: every tank of the mix within 1e-6 of its FP64 run
            cfg = dict(cfg, mg_precision=prec, reltol=1e-8, abstol=0.0, gmres_side=1 if prec else 0)
            choice = autotune.tuned(cfg, table)
            if choice:
                cfg["mg_maxgrids"] = choice["mg_maxgrids"]
                n_tuned += prec
            T = ow3d.Tank(cfg, **bath)
            autotune.apply(T, choice)
            T.set_state(E, P)
            tanks.append(T); dts.append(dt)
            dof += cfg["nx"] * cfg["ny"] * (cfg["nz"] + 1)
        t_build = time.perf_counter() - t_build
        print(f"batch {label}: {ntanks} tanks, {dof} DOF built in {t_build:.1f} s", file=sys.stderr, flush=True)
        ow3d.batch_rk4_steps(tanks, warmup, dts)
        a = time.perf_counter()
        st = ow3d.batch_rk4_steps(tanks, steps, dts)
        t_batch = time.perf_counter() - a
        print(f"batch {label}: {steps} batched steps in {t_batch:.2f} s", file=sys.stderr, flush=True)
        t_seq = 0.0
        for T, dt in zip(tanks, dts):
            a = time.perf_counter()
            T.rk4_steps(steps, 0.0, dt)
            t_seq += time.perf_counter() - a
        E_end = [T.get_state()[0] for T in tanks]
        conv = [T.rk4_steps(1, 0.0, dt) == 0 for T, dt in zip(tanks, dts)]   # one more step: did every solve converge
        res[label] = dict(dof_steps_per_s=dof * steps / t_batch, one_at_a_time_dof_steps_per_s=dof * steps / t_seq,
                          batch_gain=t_seq / t_batch, ms_per_batch_step=1e3 * t_batch / steps, status=st,
                          host_setup_s=t_build, E=E_end, conv=conv,
                          replays=sum(T.step_stats()[0] for T in tanks), captures=sum(T.step_stats()[1] for T in tanks))
        for T in tanks:
            T.close()
    errs = [float(np.max(np.abs(a - b)) / np.max(np.abs(b))) for a, b in zip(res["mixed"]["E"], res["fp64"]["E"])]
    conv = [a and b for a, b in zip(res["mixed"]["conv"], res["fp64"]["conv"])]
    err_conv = max([e for e, c in zip(errs, conv) if c], default=None)
    for r in res.values():
        del r["E"], r["conv"]
    nz = collections.Counter((d["nz"], d["order"]) for d in draws)
    return {"workload": f"{ntanks} uneven tanks (Nx, Ny in 65..513, Nz 6/9/12, order 4/6/8), seed {seed}",
            "tanks_with_tuned_schedules": n_tuned,
            "dof": sum(d["nx"] * d["ny"] * (d["nz"] + 1) for d in draws), "steps": steps, "warmup": warmup,
            "nz_order_mix": {f"{k[0]}/{k[1]}": v for k, v in sorted(nz.items())},
            "fp64": res["fp64"], "mixed": res["mixed"],
            "speedup_mixed_batched_vs_fp64_batched": res["mixed"]["dof_steps_per_s"] / res["fp64"]["dof_steps_per_s"],
            "speedup_mixed_batched_vs_fp64_one_at_a_time": res["mixed"]["dof_steps_per_s"] / res["fp64"]["one_at_a_time_dof_steps_per_s"],
            "rel_err_eta_mixed_vs_fp64_max_over_tanks": max(errs),
            "rel_err_eta_mixed_vs_fp64_max_over_converged_tanks": err_conv,
            "tanks_with_unconverged_solves": int(len(conv) - sum(conv))}


def bench_serve(seconds, rate=20.0, max_active=32, seed=0, tuned=True):
    """Request-loop workload: Poisson arrivals (rate per second) of many-tank draws with 50..500 steps and
    a flat or shoal bottom, admitted by a 200 GB HBM budget and at most max_active tanks at once, served
    for `seconds` with the FP64 and then the FP32 multigrid."""
    from ow3d import serve
    out = {"workload": f"Poisson {rate}/s tank requests, 50..500 steps, {max_active} slots, 200 GB budget, {seconds} s",
           "schedules": "tuned" if tuned else "untuned"}
    for label, prec in (("fp64", 0), ("mixed", 1)):
        reqs = serve.poisson_requests(int(rate * seconds * 2) + 8, rate, seed, mg_precision=prec, tuned=tuned)
        out[label] = serve.Server(hbm_budget=200e9, max_active=max_active).run(reqs, time_limit=seconds)
        print(f"serve {label}: {out[label]['steps']} steps, {out[label]['completed']} requests completed",
              file=sys.stderr, flush=True)
    out["speedup_mixed_vs_fp64"] = out["mixed"]["dof_steps_per_s"] / out["fp64"]["dof_steps_per_s"]
    return out


def profile_matvec(ow3d, problems, k):
    """K matvec launches on the 513x513x9 shoal tank plus one 256 MiB device-to-device copy, for
    rocprofv3 counter runs: the copy's FETCH_SIZE / WRITE_SIZE per byte calibrate the matvec's."""
    import torch
    T, cfg, dt = make_tank(ow3d, problems, 1)
    ms = T.time_matvec(k)
    src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    n = T.n
    print(json.dumps({"matvec_launches": k, "avg_ms": ms, "n": n, "ncol": T.nxg * T.nyg,
                      "algorithmic_bytes": 16 * n + 40 * T.nxg * T.nyg, "copy_bytes": 2 * src.numel() * 8}), flush=True)
    T.close()


def measure_traffic(timeout=120):
    """HBM bytes per matvec launch from two rocprofv3 counter runs of `bench.py --profile-matvec 5`
    (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each scaled by what the same counter reads for the
    256 MiB device copy the profiling run also performs.  Child processes (the profiler starts the
    program itself); None when rocprofv3 is unavailable or a run fails."""
    import csv
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3")
    if prof is None:
        return None
    res = {}
    env = dict(os.environ, TMPDIR="/tmp")
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            od = os.path.join(d, counter)
            cmd = [prof, "--pmc", counter, "-d", od, "-o", "run", "--output-format", "csv", "--",
                   sys.executable, os.path.abspath(__file__), "--profile-matvec", "5"]
            try:
                subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                               timeout=timeout)
            except subprocess.TimeoutExpired:
                return None
            path = os.path.join(od, "run_counter_collection.csv")
            if not os.path.exists(path):
                return None
            per = {}
            for r in csv.DictReader(open(path)):
                k = (int(r["Dispatch_Id"]), r["Kernel_Name"], int(r["Grid_Size"]))
                per[k] = per.get(k, 0.0) + float(r["Counter_Value"])
            mv = sorted(v for k, v in per.items() if "k_matvec3" in k[1])
            copies = [v for k, v in per.items() if "copyBuffer" in k[1] or "copy" in k[1].lower()]
            if not mv or not copies:
                return None
            copy_kib = max(copies)                         # the 256 MiB device copy
            expect_kib = 256.0 * 1024                      # it reads 256 MiB and writes 256 MiB
            res[counter] = {"kib_per_launch": mv[len(mv) // 2], "copy_kib": copy_kib, "scale": expect_kib / copy_kib}
    fetch = res["FETCH_SIZE"]["kib_per_launch"] * res["FETCH_SIZE"]["scale"] * 1024
    write = res["WRITE_SIZE"]["kib_per_launch"] * res["WRITE_SIZE"]["scale"] * 1024
    return {"bytes": fetch + write, "fetch_bytes": fetch, "write_bytes": write,
            "fetch_scale": res["FETCH_SIZE"]["scale"], "write_scale": res["WRITE_SIZE"]["scale"]}


def make_tank_state(problems):
    cfg, bath, E, P, dt, zones = problems.workload(WORKLOAD)
    return E, P


if __name__ == "__main__":
    main()
