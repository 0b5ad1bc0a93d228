"""Request loop for many wave-tank simulations on one GPU.

Requests (a tank, a step count, a precision) arrive over time.  A request is admitted as soon as it has
arrived, a slot is free (at most `max_active` tanks on the device) and its device memory
(ow3d.estimate_bytes) fits the remaining HBM budget.  An admitted tank keeps its state on the device
and advances one RK4 step at a time on its own stream: the loop enqueues a tank's next step as soon as
its previous one has settled, so steps of different tanks overlap on the GPU and tanks join and leave
between steps (continuous batching, no barrier between tanks).  Tanks are built by a worker thread
(host set-up of grids, stencils and the multigrid hierarchy) while the loop keeps the GPU busy.

Reported: aggregate DOF-steps/s (DOF = Nx*Ny*(Nz+1) per tank, as OceanWave3D.f90:29), p50/p99 of the
per-step latency (from the moment a step is enqueued to the moment the loop sees it settled), queueing
delay, admissions and completions, and how many completed requests had a Laplace solve that missed its
tolerance (the reference ignores this, iterative_solution.f90:115-134).  A request whose tank can never
fit the HBM budget is rejected on arrival instead of blocking the queue.  Each request's result does not
depend on what else runs: tanks share nothing on the device, so a request's final state is bit-identical
alone and in the mix.

    python -m ow3d.serve --rate 20 --duration 20 --max-active 32 --precision both
"""
import argparse
import collections
import json
import queue
import threading
import time

import numpy as np

import ow3d
from ow3d import autotune, problems

PRECISIONS = {"fp64": 0, "mixed": 1}
_TABLE = autotune.load()


class Request:
    """One tank simulation: a problems.batch_draw(...) dict (the tank's fields are built when it is
    admitted), nsteps RK4 steps, multigrid precision (0 FP64, 1 FP32); tuned: use the autotuner's cached
    schedule for the tank's class (else the default levels and launch choices)."""

    def __init__(self, rid, draw, nsteps, mg_precision, arrival=0.0, tuned=True):
        self.rid = rid
        self.draw = draw
        self.nsteps = int(nsteps)
        self.mg_precision = int(mg_precision)
        self.arrival = float(arrival)
        cfg = problems.batch_cfg(draw)
        # the FP32 multigrid runs with flexible GMRES, so that every tank of the varied mix stays within 1e-6 of
        # its FP64 run (left preconditioning with an FP32 V cycle does not on some tanks)
        cfg.update(mg_precision=self.mg_precision, reltol=1e-8, abstol=0.0,
                   gmres_side=1 if self.mg_precision >= 1 else 0)
        self.choice = autotune.tuned(cfg, _TABLE) if tuned else None   # the class's cached schedule
        if self.choice:
            cfg["mg_maxgrids"] = self.choice["mg_maxgrids"]
        self.cfg = cfg
        self.dof = cfg["nx"] * cfg["ny"] * (cfg["nz"] + 1)
        self.bytes = ow3d.estimate_bytes(cfg)


def poisson_requests(n, rate, seed=0, mg_precision=1, steps=(50, 500), sizes=problems.BATCH_SIZES, tuned=True):
    """n requests with exponential inter-arrival times (mean 1/rate s), tanks drawn as the many-tank
    workload (flat or Gaussian-shoal bottom at random), step counts uniform in [steps[0], steps[1]]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += float(rng.exponential(1.0 / rate))
        bottom = "shoal" if rng.uniform() < 0.5 else "flat"
        draw = problems.batch_draw(int(seed) * 1000003 + i, bottom=bottom, sizes=sizes)
        out.append(Request(i, draw, int(rng.integers(steps[0], steps[1] + 1)), mg_precision, t, tuned))
    return out


def build_tank(req):
    """The tank of a request in its initial state (the same set-up alone and in the mix)."""
    cfg, bath, E, P, dt, zones = problems.batch_tank_from(req.draw)
    T = ow3d.Tank(req.cfg, **bath)
    autotune.apply(T, req.choice)
    for z in zones:
        T.add_relax_zone(*z)
    T.set_state(E, P)
    return T, dt


def run_alone(req):
    """Final (E, P, W, phi) of a request run by itself."""
    T, dt = build_tank(req)
    T.rk4_steps(req.nsteps, 0.0, dt)
    out = T.get_state()
    T.close()
    return out


class _Active:
    __slots__ = ("req", "tank", "dt", "done", "submitted", "admitted", "unconverged")

    def __init__(self, req, tank, dt, now):
        self.req, self.tank, self.dt = req, tank, dt
        self.done = 0
        self.submitted = now
        self.admitted = now
        self.unconverged = 0   # steps in which a solve missed its tolerance


class Server:
    """Continuous-batching request loop (see the module docstring)."""

    def __init__(self, hbm_budget=200e9, max_active=32, keep_results=False, poll_s=50e-6):
        self.hbm_budget = float(hbm_budget)
        self.max_active = int(max_active)
        self.keep_results = keep_results
        self.poll_s = poll_s

    def run(self, requests, time_limit=None):
        """Serves `requests` (arrival times in seconds from the start).  Stops when all are done or, with
        time_limit, when that many seconds have passed (requests still running are then dropped and only
        completed steps count).  Returns a summary dict (results by request id when keep_results)."""
        waiting = collections.deque(sorted(requests, key=lambda r: r.arrival))
        built = queue.Queue()
        want = queue.Queue()
        setup = []   # host set-up seconds of each built tank
        stop = threading.Event()

        def builder():   # host set-up of admitted requests, off the loop's thread
            while not stop.is_set():
                try:
                    req = want.get(timeout=0.01)
                except queue.Empty:
                    continue
                try:
                    a = time.perf_counter()
                    tk = build_tank(req)
                    setup.append(time.perf_counter() - a)
                    built.put((req, tk, None))
                except Exception as e:   # reported by the loop
                    built.put((req, None, e))

        worker = threading.Thread(target=builder, daemon=True)
        worker.start()
        active = {}
        reserved = 0.0
        in_build = 0
        lat, qdelay, results = [], [], {}
        steps_done = dof_steps = completed = 0
        rejected = []                      # request ids larger than the whole budget
        unconverged_requests = unconverged_steps = 0
        t0 = time.perf_counter()
        try:
            while waiting or active or in_build:
                now = time.perf_counter() - t0
                if time_limit is not None and now > time_limit:
                    break
                # a request that cannot fit even an empty budget is rejected when it arrives
                while waiting and waiting[0].arrival <= now and waiting[0].bytes > self.hbm_budget:
                    rejected.append(waiting.popleft().rid)
                # admission: arrived, a free slot, memory within the budget
                while (waiting and waiting[0].arrival <= now and len(active) + in_build < self.max_active
                       and reserved + waiting[0].bytes <= self.hbm_budget):
                    req = waiting.popleft()
                    reserved += req.bytes
                    in_build += 1
                    qdelay.append(now - req.arrival)
                    want.put(req)
                while True:
                    try:
                        req, tk, e = built.get_nowait()
                    except queue.Empty:
                        break
                    in_build -= 1
                    if e is not None:
                        raise e
                    T, dt = tk
                    a = _Active(req, T, dt, time.perf_counter() - t0)
                    T.rk4_steps_async(1, dt)
                    active[req.rid] = a
                progressed = False
                for rid in list(active):
                    a = active[rid]
                    if not a.tank.query():
                        continue
                    progressed = True
                    t = time.perf_counter() - t0
                    lat.append(t - a.submitted)
                    a.done += 1
                    a.unconverged += a.tank.steps_status()
                    steps_done += 1
                    dof_steps += a.req.dof
                    if a.done < a.req.nsteps:
                        a.submitted = t
                        a.tank.rk4_steps_async(1, a.dt)
                        continue
                    if self.keep_results:
                        results[rid] = a.tank.get_state()
                    a.tank.close()
                    reserved -= a.req.bytes
                    completed += 1
                    unconverged_steps += a.unconverged
                    unconverged_requests += a.unconverged > 0
                    del active[rid]
                if not progressed:
                    time.sleep(self.poll_s)
        finally:
            stop.set()
            worker.join()
            for a in active.values():
                a.tank.wait()
                a.tank.close()
            while not built.empty():
                req, tk, e = built.get_nowait()
                if tk is not None:
                    tk[0].close()
        wall = time.perf_counter() - t0
        lat_ms = 1e3 * np.asarray(lat) if lat else np.zeros(1)
        out = {
            "wall_s": wall,
            "requests": len(requests),
            "admitted": len(qdelay),
            "completed": completed,
            "rejected": rejected,
            "completed_with_unconverged_solves": unconverged_requests,
            "unconverged_steps_of_completed": unconverged_steps,
            "steps": steps_done,
            "dof_steps_per_s": dof_steps / wall if wall > 0 else 0.0,
            "step_latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99)),
                                "mean": float(lat_ms.mean())},
            "queue_delay_s": {"p50": float(np.percentile(qdelay, 50)) if qdelay else 0.0,
                              "p99": float(np.percentile(qdelay, 99)) if qdelay else 0.0},
            "max_active": self.max_active,
            "hbm_budget_GB": self.hbm_budget / 1e9,
            "setup_s": {"tanks": len(setup), "p50": float(np.percentile(setup, 50)) if setup else 0.0,
                        "max": float(max(setup)) if setup else 0.0, "total": float(sum(setup))},
        }
        if self.keep_results:
            out["results"] = results
        return out


def main(argv=None):
    ap = argparse.ArgumentParser(description="Serve a Poisson stream of wave-tank simulations on one GPU.")
    ap.add_argument("--rate", type=float, default=20.0, help="mean arrivals per second")
    ap.add_argument("--requests", type=int, default=400)
    ap.add_argument("--duration", type=float, default=20.0, help="seconds to serve (completed steps count)")
    ap.add_argument("--max-active", type=int, default=32)
    ap.add_argument("--hbm-budget-gb", type=float, default=200.0)
    ap.add_argument("--precision", choices=["fp64", "mixed", "both"], default="both")
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args(argv)
    out = {}
    for name in (["fp64", "mixed"] if args.precision == "both" else [args.precision]):
        reqs = poisson_requests(args.requests, args.rate, args.seed, mg_precision=PRECISIONS[name])
        srv = Server(hbm_budget=args.hbm_budget_gb * 1e9, max_active=args.max_active)
        out[name] = srv.run(reqs, time_limit=args.duration)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
