"""The stream-function wave as the second, analytic oracle of the time-stepping path.

* The solver (ow3d/streamfunction.py, Fenton's method) is pinned by the linear dispersion relation, the
  deep-water Stokes speed c = c0 (1 + (ka)^2 / 2), and the free-surface conditions holding between its
  collocation nodes.
* A steep wave (kh = 1, H/L = 0.05, the CPU workload's wave) of permanent form must travel as
  eta(x - c t): an x-z tank of 8 wavelengths with 1-wavelength damping zones at the walls, 32 points per
  wavelength, Nz = 9, Courant 0.5, one wave period; compared on the middle two wavelengths, which no wall
  or zone disturbance reaches within a period.  The run starts from a consistent W (one Laplace solve at
  t = 0); the reference starts its first stage with W = 0 (Runge_Kutta_4.f90), which costs a first-order
  error in dt and is why the tests solve first.
  Bound: |eta - eta_SF| <= 3e-4 H on the oracle and on the GPU; GPU and oracle trajectories agree to 1e-9.
"""
import math

import numpy as np
import pytest

from conftest import oracle_params

NLAM, PPW = 8, 32


def _sf():
    from ow3d import streamfunction
    return streamfunction


def test_stream_function_limits():
    sf = _sf()
    w = sf.StreamFunctionWave(1.0, 1.0, 1e-4, N=8)
    assert abs(w.c / math.sqrt(9.81 * math.tanh(1.0)) - 1.0) < 1e-7
    k, a = 1.0, 0.05
    deep = sf.StreamFunctionWave(k, 20.0, 2 * a, N=24)
    stokes = 1.0 + (k * a) ** 2 / 2
    assert abs(deep.c / math.sqrt(9.81 / k) - stokes) < 2e-5 * stokes


def test_stream_function_conditions_between_nodes():
    sf = _sf()
    w = sf.StreamFunctionWave(1.0, 1.0, 0.05 * 2 * math.pi, N=32)
    assert abs((w.eta(0.0) - w.eta(math.pi)) - w.H) < 1e-12
    X = np.linspace(0.0, 2 * math.pi, 257)
    e = w.eta(X)
    j = np.arange(1, w.N + 1)
    kz = w.k * (e[:, None] + w.d)
    ch = np.cosh(j * kz) / np.cosh(j * w.k * w.d)
    sh = np.sinh(j * kz) / np.cosh(j * w.k * w.d)
    u = -w.ubar + (j * w.k * ch * np.cos(j * w.k * X[:, None])) @ w.B
    v = (j * w.k * sh * np.sin(j * w.k * X[:, None])) @ w.B
    assert np.max(np.abs(0.5 * (u * u + v * v) + 9.81 * e - w.R)) < 1e-10
    psi = -w.ubar * (e + w.d) + (sh * np.cos(j * w.k * X[:, None])) @ w.B
    assert np.max(np.abs(psi + w.Q)) < 1e-10
    # the surface potential is phi(x, eta(x)) of the same series
    E, P = w.surface(X)
    assert np.allclose(E, e) and np.allclose(P, w.potential(X, e))


def sf_case():
    from ow3d import problems
    sf = _sf()
    nx = NLAM * PPW + 1
    wave, x, E, P, dt = sf.sf_tank(nx, NLAM)
    lam = wave.wavelength
    nsteps = int(round(wave.period / dt))
    dt = wave.period / nsteps
    bath = problems.flat_bottom(nx, 1, NLAM * lam, 1.0)
    zones = problems.sponge_zones(nx, 1, NLAM * lam, 1.0, lam)
    p = oracle_params(nx, 1, 9, lx=NLAM * lam, ly=1.0, reltol=1e-12, gmres_m=31)
    return wave, x, E, P, dt, nsteps, bath, zones, p


def consistent_start(T, E, P):
    """Ghost update, one Laplace solve with the surface potential as Dirichlet data, W from it."""
    E2, P2, _ = T.surface(E, P) if hasattr(T, "surface") else T.set_surface(E, P)
    b = np.zeros(T.shape3)
    b[0, 1:-1, -1] = P2[0, 1:-1]
    phi = T.solve(b, np.zeros(T.shape3))[0]
    return E2, P2, T.vertical_velocity(phi), phi


def middle_error(wave, x, E, t):
    lam = wave.wavelength
    mid = (x > 3 * lam) & (x < 5 * lam)
    return float(np.max(np.abs(E[0, mid] - wave.eta(x, t)[mid])) / wave.H)


def test_oracle_translates_stream_function_wave(oracle_mod):
    wave, x, E, P, dt, nsteps, bath, zones, p = sf_case()
    T = oracle_mod.OracleTank(p, bath["h"], bath["hx"], bath["hy"], bath["hxx"], bath["hyy"])
    for z in zones:
        T.add_relax_zone(*z)
    T.set_state(*consistent_start(T, E, P))
    for s in range(nsteps):
        T.rk4_step(s * dt, dt)
    err = middle_error(wave, x, T.get_state()[0], nsteps * dt)
    assert err < 3e-4, err
    T.close()


@pytest.mark.gpu
def test_gpu_translates_stream_function_wave(oracle_mod):
    import ow3d
    wave, x, E, P, dt, nsteps, bath, zones, p = sf_case()
    cfg = {k: v for k, v in p.items() if k not in ("mg_mode", "cd_quirk")}
    cfg["mg_precision"] = 0
    G = ow3d.Tank(cfg, bath["h"], bath["hx"], bath["hy"], bath["hxx"], bath["hyy"])
    O = oracle_mod.OracleTank(p, bath["h"], bath["hx"], bath["hy"], bath["hxx"], bath["hyy"])
    for z in zones:
        G.add_relax_zone(*z)
        O.add_relax_zone(*z)
    G.set_state(*consistent_start(G, E, P))
    O.set_state(*consistent_start(O, E, P))
    G.rk4_steps(nsteps, 0.0, dt)
    for s in range(nsteps):
        O.rk4_step(s * dt, dt)
    Eg, Pg, _, _ = G.get_state()
    Eo, Po, _, _ = O.get_state()
    assert middle_error(wave, x, Eg, nsteps * dt) < 3e-4
    assert np.max(np.abs(Eg - Eo)) <= 1e-9 * np.max(np.abs(Eo))
    assert np.max(np.abs(Pg - Po)) <= 1e-9 * np.max(np.abs(Po))
    G.close()
    O.close()
