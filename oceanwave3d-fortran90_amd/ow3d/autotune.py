"""Autotuner of per-tank launch schedules.

For a class of tanks (Nx, Ny, Nz, stencil order) it measures, on the GPU, RK4 steps of a representative
tank under each candidate and keeps the fastest:

* the V-cycle schedule: the number of multigrid levels (MGmaxgrids), from the full hierarchy down to the
  fewest levels whose coarsest level the dense solve accepts (at most 4096 unknowns, ow3d.mg_plan);
  fewer levels trade launch-latency-bound small levels for a larger dense coarse solve;
* the matvec kernel (ow3d_tune: column kernel with the widest tile, with 16 x 8 tiles, with each column
  split over two waves, or the point kernel), timed with ow3d_time_matvec;
* the block size of the zebra sweeps (automatic, 256 or 1024 threads).

None of these changes results beyond the multigrid depth, which is the reference's own MGmaxgrids
input (the solve still runs to its tolerance).  Choices are cached by class in tuning.json next to this
file; tuned(cfg) returns the cached choice or None.

    python -m ow3d.autotune --batch 64 --seed 0      # tune the classes of the 64-tank workload
    python -m ow3d.autotune --batch 64 --workload shoal513 --matvec-only   # re-time the matvec choice only
"""
import argparse
import json
import os
import time

import ow3d
from ow3d import problems

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
MAX_COARSE = 4096


def class_key(cfg):
    """Nx x Ny x Nz / stencil order / multigrid precision."""
    return f"{cfg['nx']}x{cfg['ny']}x{cfg['nz']}/o{2 * cfg.get('alpha', 3)}/p{cfg.get('mg_precision', 0)}"


def load(path=CACHE):
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def save(table, path=CACHE):
    with open(path + ".tmp", "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(path + ".tmp", path)


def tuned(cfg, table=None):
    """Cached choice for cfg's class: dict(mg_maxgrids, matvec_variant, sweep_threads) or None."""
    table = load() if table is None else table
    return table.get(class_key(cfg))


def level_candidates(cfg, keep=3):
    """Level counts to try: the fewest levels with an accepted coarsest level, then up to keep - 1 more."""
    full, _ = ow3d.mg_plan(dict(cfg, mg_maxgrids=0))
    ok = []
    for L in range(2, full + 1):
        n, nc = ow3d.mg_plan(dict(cfg, mg_maxgrids=L))
        if nc <= MAX_COARSE:
            ok.append(n)
    ok = sorted(set(ok))
    return ok[:keep] if ok else [full]


def _time_steps(T, dt, warmup, steps):
    T.rk4_steps(warmup, 0.0, dt)
    a = time.perf_counter()
    T.rk4_steps(steps, 0.0, dt)
    return (time.perf_counter() - a) / steps


def tune_class(spec, mg_precision=1, warmup=3, steps=3, log=None):
    """Tunes the class of spec = (cfg, bathymetry, E, P, dt, zones) and returns the choice with timings."""
    cfg0, bath, E, P, dt, zones = spec
    base = dict(cfg0, mg_precision=mg_precision, gmres_m=23, reltol=1e-8, abstol=0.0)
    for k in ("mg_maxgrids",):
        base.pop(k, None)

    def make(levels):
        T = ow3d.Tank(dict(base, mg_maxgrids=levels), **bath)
        for z in zones:
            T.add_relax_zone(*z)
        T.set_state(E, P)
        return T

    best = None
    per_levels = {}
    for L in level_candidates(base):
        T = make(L)
        ms = 1e3 * _time_steps(T, dt, warmup, steps)
        per_levels[L] = ms
        if best is None or ms < best[1]:
            if best is not None:
                best[0].close()
            best = (T, ms, L)
        else:
            T.close()
    T, _, L = best
    mv = {}
    for v in ((0, 1, 2, 3) if base["ny"] > 1 else (0,)):
        T.tune(v, 0)
        T.time_matvec(3)
        mv[v] = 1e3 * T.time_matvec(30)
    v_best = min(mv, key=mv.get)
    thr = {}
    for n in (0, 256, 1024):
        T.tune(v_best, n)
        thr[n] = 1e3 * _time_steps(T, dt, 1, steps)
    n_best = min(thr, key=thr.get)
    T.close()
    choice = {"mg_maxgrids": L, "matvec_variant": v_best, "sweep_threads": n_best,
              "ms_per_step": round(thr[n_best], 3), "levels_ms": {str(k): round(v, 3) for k, v in per_levels.items()},
              "matvec_us": {str(k): round(v, 2) for k, v in mv.items()},
              "sweep_threads_ms": {str(k): round(v, 3) for k, v in thr.items()}}
    if log:
        log(f"{class_key(base)}: {choice}")
    return choice


def retune_matvec(spec, choice, mg_precision=1):
    """Re-times only the matvec kernels of a class whose schedule is cached (after a kernel change) and
    returns the choice with the fastest variant."""
    cfg0, bath, E, P, dt, zones = spec
    base = dict(cfg0, mg_precision=mg_precision, gmres_m=23, reltol=1e-8, abstol=0.0, mg_maxgrids=choice["mg_maxgrids"])
    T = ow3d.Tank(base, **bath)
    T.set_state(E, P)
    mv = {}
    for v in ((0, 1, 2, 3) if base["ny"] > 1 else (0,)):
        T.tune(v, choice["sweep_threads"])
        T.time_matvec(3)
        mv[v] = 1e3 * T.time_matvec(30)
    T.close()
    out = dict(choice)
    out["matvec_variant"] = min(mv, key=mv.get)
    out["matvec_us"] = {str(k): round(v, 2) for k, v in mv.items()}
    return out


def apply(T, choice):
    """Applies a cached choice's launch knobs to a tank built with its mg_maxgrids."""
    if choice:
        T.tune(choice["matvec_variant"], choice["sweep_threads"])


def main(argv=None):
    ap = argparse.ArgumentParser(description="Tune the tank classes of the many-tank workload.")
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--retune", action="store_true")
    ap.add_argument("--cache", default=CACHE, help="table to read and extend")
    ap.add_argument("--workload", action="append", default=[], help="also tune a problems.workload tank (e.g. shoal513)")
    ap.add_argument("--precisions", type=int, nargs="+", default=[0, 1])
    ap.add_argument("--matvec-only", action="store_true", help="re-time only the matvec of the cached classes")
    args = ap.parse_args(argv)
    table = load(args.cache)
    specs = [(lambda n=n: problems.workload(n)) for n in args.workload]
    specs += [(lambda d=problems.batch_draw(args.seed * 1000003 + i): problems.batch_tank_from(d)) for i in range(args.batch)]
    cfgs = [problems.workload(n)[0] for n in args.workload]
    cfgs += [problems.batch_cfg(problems.batch_draw(args.seed * 1000003 + i)) for i in range(args.batch)]
    retimed = set()
    for prec in args.precisions:
        for make_spec, cfg in zip(specs, cfgs):
            key = class_key(dict(cfg, mg_precision=prec))
            if args.matvec_only:
                if key in table and key not in retimed:
                    retimed.add(key)
                    table[key] = retune_matvec(make_spec(), table[key], mg_precision=prec)
                    print(f"{key}: matvec {table[key]['matvec_variant']} {table[key]['matvec_us']}", flush=True)
                    save(table, args.cache)
                continue
            if key in table and not args.retune:
                continue
            table[key] = tune_class(make_spec(), mg_precision=prec, log=lambda s: print(s, flush=True))
            save(table, args.cache)
    print(json.dumps({"classes": len(table)}))


if __name__ == "__main__":
    main()
