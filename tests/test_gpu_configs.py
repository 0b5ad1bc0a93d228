"""The benchmark configurations on the GPU, at the sizes they name or as close as the oracle allows:

* flat-bottom 3D JONSWAP tank (257 x 257 x 9, 6th order, 64 Airy components, Hs = 0.02 h, k_p h = 1.5):
  a 65 x 65 x 9 tank of the same family (same resolution, 16 points per peak wavelength) stepped against
  the oracle, and the full 257 x 257 x 9 tank stepped 20 times with every solve converging;
* many uneven tanks: a batch holding 257^2 and 513^2 tanks next to small ones, each bit-identical to its
  run alone;
* serving: a request loop whose tanks are drawn from the many-tank sizes (65 .. 513), each result
  bit-identical to its run alone.
"""
import math

import numpy as np
import pytest

from conftest import oracle_params

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


def jonswap_tank(n, seed=0):
    """Flat-bottom JONSWAP tank of the 257^2 workload's family on an n x n grid (16 points per peak
    wavelength, dt = 0.5 dx / c_p)."""
    from ow3d import problems
    lam = 2 * math.pi / 1.5
    L = (n - 1) / 16 * lam
    bath = problems.flat_bottom(n, n, L, L)
    E, P, info = problems.random_waves(n, n, L, L, hs=0.02, kph=1.5, seed=seed)
    dt = 0.5 * (L / (n - 1)) / info["cp"]
    return dict(nx=n, ny=n, nz=9, lx=L, ly=L), bath, E, P, dt


def test_jonswap_tank_matches_the_oracle(oracle_mod):
    """65 x 65 x 9 flat JONSWAP tank, 20 RK4 steps, solves driven to rtol 1e-13 on both sides."""
    import ow3d
    cfg, bath, E, P, dt = jonswap_tank(65, seed=3)
    p = oracle_params(cfg["nx"], cfg["ny"], cfg["nz"], lx=cfg["lx"], ly=cfg["ly"], reltol=1e-13)
    T_o = oracle_mod.OracleTank(p, bath["h"], bath["hx"], bath["hy"], bath["hxx"], bath["hyy"])
    T_g = ow3d.Tank({k: v for k, v in p.items() if k not in ("mg_mode", "cd_quirk")}, **bath)
    T_o.set_state(E, P)
    T_g.set_state(E, P)
    nsteps = 20
    for s in range(nsteps):
        T_o.rk4_step(s * dt, dt)
    T_g.rk4_steps(nsteps, 0.0, dt)
    Eo, Po, Wo, _ = T_o.get_state()
    Eg, Pg, Wg, _ = T_g.get_state()
    assert rel(Eg, Eo) < 1e-10, rel(Eg, Eo)
    assert rel(Pg, Po) < 1e-10, rel(Pg, Po)
    assert rel(Wg, Wo) < 1e-9, rel(Wg, Wo)


@pytest.mark.parametrize("prec", [0, 1])
def test_jonswap_tank_at_257_converges_every_solve(prec):
    """The 257 x 257 x 9 flat JONSWAP tank itself, 20 RK4 steps at GMRES(23) rtol 1e-8: every Laplace
    solve converges, the state stays finite and the wave height stays that of a linear sea (within 20 %
    of the initial one).  The FP32 multigrid run is then within 1e-6 of the FP64 one."""
    import ow3d
    from ow3d import problems
    cfg, bath, E, P, dt, zones = problems.workload("tank257", seed=0)
    out = {}
    for q in sorted({0, prec}):
        T = ow3d.Tank(dict(cfg, mg_precision=q, reltol=1e-8, gmres_m=23), **bath)
        T.set_state(E, P)
        st = [T.rk4_steps(1, s * dt, dt) for s in range(20)]
        assert st == [0] * 20, st
        tot, solves, mn, mx = T.stats()
        assert solves == 80 and mx < 23
        out[q] = T.get_state()
        T.close()
    En = out[prec][0]
    assert all(np.all(np.isfinite(a)) for a in out[prec])
    h0 = 4 * np.std(E[1:-1, 1:-1])
    assert abs(4 * np.std(En[1:-1, 1:-1]) / h0 - 1.0) < 0.2
    if prec:
        assert rel(out[1][0], out[0][0]) < 1e-6 and rel(out[1][1], out[0][1]) < 1e-6


def _batch_tank(seed, size, prec):
    import ow3d
    from ow3d import problems
    cfg, bath, E, P, dt, zones = problems.batch_tank(seed, sizes=(size,))
    T = ow3d.Tank(dict(cfg, mg_precision=prec, reltol=1e-8, gmres_m=23), **bath)
    T.set_state(E, P)
    return T, dt


def test_batch_with_257_and_513_tanks_is_bit_identical_to_each_tank_alone():
    import ow3d
    specs = [(101, 513), (102, 257), (103, 65), (104, 129)]
    nsteps = 3
    alone = []
    for seed, size in specs:
        print(f"alone: seed {seed}, {size}^2", flush=True)
        T, dt = _batch_tank(seed, size, 1)
        st = T.rk4_steps(nsteps, 0.0, dt)
        assert st in (0, 1)
        alone.append(T.get_state())
        T.close()
    tanks = [_batch_tank(seed, size, 1) for seed, size in specs]
    print("batch:", [(T.nxg - 2, T.nyg - 2, T.nzg - 1) for T, _ in tanks], flush=True)
    assert max(T.nxg for T, _ in tanks) == 515
    ow3d.batch_rk4_steps([T for T, _ in tanks], nsteps, [dt for _, dt in tanks])
    for (T, _), ref in zip(tanks, alone):
        assert same(T.get_state(), ref)
        T.close()


def test_request_loop_over_the_many_tank_sizes_equals_each_request_alone():
    """Four requests drawn from sizes 65 .. 513 (one of them 513 x 513 x 9), two slots: each result equals
    its run alone, bit for bit."""
    from ow3d import serve
    reqs = serve.poisson_requests(4, rate=100.0, seed=18, mg_precision=1, steps=(2, 4))
    assert any(r.cfg["nx"] == r.cfg["ny"] == 513 for r in reqs), [(r.cfg["nx"], r.cfg["ny"]) for r in reqs]
    out = serve.Server(hbm_budget=200e9, max_active=2, keep_results=True).run(reqs)
    assert out["completed"] == 4 and out["steps"] == sum(r.nsteps for r in reqs)
    for r in reqs:
        assert same(out["results"][r.rid], serve.run_alone(r)), r.rid
