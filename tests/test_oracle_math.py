"""Mathematical pins of the oracle for what the reference's own code cannot check here
(SPARSKIT2 GMRES and HSL MA41 are not in the reference tree, so iterative_solution, dc and the
multigrid driver cannot be built) and independent checks of the rest:

* Fornberg weights: closed-form centred 6th-order weights; exactness on polynomials.
* High-order operator: 6th-order convergence to zero on harmonic functions in physical space mapped
  through the sigma transform (flat and shoaling bottom, moving surface), which exercises every
  interior coefficient, the kinematic bottom row and the wall Neumann rows.
* Zebra-line multigrid operator = the reference-assembled level matrix without the cross-derivative
  replication; V-cycle contraction.
* GMRES / defect correction: convergence to a dense direct solve of the assembled system.
* Laplace solve + vertical velocity: the linear standing-wave Dirichlet-to-Neumann known answer
  (TestCodeDucrozet2.f90:364-455 with linearstandingwave2D.f90:23-31).
* RK4 driver: linear standing wave over a period against the analytic solution.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import oracle_params

import oracle


def P():
    import ow3d.problems as pr
    return pr


def test_fornberg_closed_forms():
    x = np.arange(-3, 4, dtype=float)
    c = oracle.fornberg(0.0, x, 2)
    assert np.allclose(c[:, 1] * 60, [-1, 9, -45, 0, 45, -9, 1], atol=1e-12)
    assert np.allclose(c[:, 2] * 180, [2, -27, 270, -490, 270, -27, 2], atol=1e-11)
    rng = np.random.default_rng(0)
    nodes = np.sort(rng.uniform(-1, 1, 7))
    z0 = 0.13
    w = oracle.fornberg(z0, nodes, 2)
    for p in range(7):
        f = nodes ** p
        d1 = p * z0 ** (p - 1) if p >= 1 else 0.0
        d2 = p * (p - 1) * z0 ** (p - 2) if p >= 2 else 0.0
        assert abs(w[:, 1] @ f - d1) < 1e-9 and abs(w[:, 2] @ f - d2) < 1e-8


def _harmonic_residuals(nx, ny, nz, L, shoal, eta_amp):
    pr = P()
    bath = (pr.gaussian_shoal(nx, ny, L, L, amp=0.3, width=0.25) if shoal else pr.flat_bottom(nx, ny, L, L))
    p = oracle_params(nx, ny, nz, lx=L, ly=L, precond=0)
    T = oracle.OracleTank(p, bath["h"], bath["hx"], bath["hy"], bath["hxx"], bath["hyy"])
    x, y, s = T.grid()
    X, Y = np.meshgrid(x, y)
    E = eta_amp * np.cos(2 * np.pi * X / L) * (np.cos(2 * np.pi * Y / L) if ny > 1 else 1.0)
    E, _, _ = T.surface(E, np.zeros_like(E))
    # physical harmonic function  f = exp(k z) cos(kx x) cos(ky y),  k = |(kx, ky)|
    kx = 2 * np.pi / L
    ky = (2 * np.pi / L) if ny > 1 else 0.0
    k = math.hypot(kx, ky)
    h = bath["h"]
    d = E + h
    sig = s.copy()
    sig[0] = s[0]  # physical position of the ghost level follows z = sigma d - h as well
    Z = sig[None, None, :] * d[:, :, None] - h[:, :, None]
    cx = np.cos(kx * X)[:, :, None]
    cy = np.cos(ky * Y)[:, :, None]
    phi = np.exp(k * Z) * cx * cy
    r = T.matvec(phi)
    jj = slice(1, -1) if ny > 1 else slice(0, 1)
    interior = np.abs(r[jj, 1:-1, 1:-1]).max()
    # bottom row = d phi/dz + h_x d phi/dx + h_y d phi/dy at z = -h (physical)
    zb = -h
    fz = k * np.exp(k * zb) * cx[:, :, 0] * cy[:, :, 0]
    fx = -kx * np.exp(k * zb) * np.sin(kx * X) * cy[:, :, 0]
    fy = -ky * np.exp(k * zb) * cx[:, :, 0] * np.sin(ky * Y)
    bot = fz + bath["hx"] * fx + bath["hy"] * fy
    bottom = np.abs(r[jj, 1:-1, 0] - bot[jj, 1:-1]).max()
    return interior, bottom


@pytest.mark.parametrize("ny", [1, 33])
@pytest.mark.parametrize("shoal", [False, True])
def test_highorder_operator_converges_on_harmonic_functions(ny, shoal):
    errs = []
    for n in (17, 33, 65):
        nyy = 1 if ny == 1 else n
        errs.append(_harmonic_residuals(n, nyy, 17, 4.0, shoal, 0.02))
    e_int = [e[0] for e in errs]
    e_bot = [e[1] for e in errs]
    # horizontal refinement at fixed vertical resolution: error drops by >= 2^3 per halving until
    # the vertical error floor is reached
    assert e_int[2] < e_int[0] / 30.0, e_int
    assert e_bot[2] < e_bot[0] / 10.0 or e_bot[2] < 1e-6, e_bot   # flat bottom: vertical error floor only


def test_zebra_operator_equals_reference_assembly_without_replication():
    pr = P()
    nx = ny = 17
    bath = pr.gaussian_shoal(nx, ny, 10.0, 10.0)
    args = (bath["h"], bath["hx"], bath["hy"], bath["hxx"], bath["hyy"])
    T0 = oracle.OracleTank(oracle_params(nx, ny, 9, lx=10.0, ly=10.0, mg_mode=0, cd_quirk=0), *args)
    T1 = oracle.OracleTank(oracle_params(nx, ny, 9, lx=10.0, ly=10.0, mg_mode=1), *args)
    rng = np.random.default_rng(1)
    for level in range(T0.nlev):
        d = T0.level_dims(level)
        v = rng.standard_normal((d[1], d[0], d[2]))
        a = T0.lowmatvec(level, v)
        b = T1.lowmatvec(level, v)
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(a))


@pytest.mark.parametrize("mode", [0, 1])
def test_vcycle_contracts(mode):
    pr = P()
    nx = ny = 33
    bath = pr.flat_bottom(nx, ny, 10.0, 10.0)
    T = oracle.OracleTank(oracle_params(nx, ny, 9, lx=10.0, ly=10.0, mg_mode=mode, cd_quirk=mode == 0),
                          bath["h"], bath["hx"], bath["hy"], bath["hxx"], bath["hyy"])
    top = T.nlev - 1
    b = np.random.default_rng(2).standard_normal(T.shape3)
    x = np.zeros(T.shape3)
    res = []
    for _ in range(8):
        r = b - T.lowmatvec(top, x)
        res.append(np.linalg.norm(r))
        x = x + T.precond(r)
    rate = (res[-1] / res[-4]) ** (1 / 3)
    assert rate < (0.6 if mode == 0 else 0.35), rate


@pytest.mark.parametrize("solver,side", [(1, 0), (1, 1), (0, 0)])
def test_krylov_converges_to_dense_solution(solver, side):
    """GMRES with left preconditioning (the reference's) and flexible right preconditioning, and defect
    correction, against numpy's dense solve of the assembled high-order system."""
    pr = P()
    nx, ny, nz = 9, 9, 5
    bath = pr.gaussian_shoal(nx, ny, 4.0, 4.0)
    p = oracle_params(nx, ny, nz, alpha=2, beta=2, gamma=2, lx=4.0, ly=4.0, solver=solver, reltol=1e-14,
                      gmres_m=31, gmres_side=side)
    T = oracle.OracleTank(p, bath["h"], bath["hx"], bath["hy"], bath["hxx"], bath["hyy"])
    n = T.n
    A = np.zeros((n, n))
    e = np.zeros(T.shape3)
    for c in range(n):
        e.flat[c] = 1.0
        A[:, c] = T.matvec(e).ravel()
        e.flat[c] = 0.0
    b = np.zeros(T.shape3)
    b[1:-1, 1:-1, -1] = np.random.default_rng(3).standard_normal((ny, nx))
    x_ref = np.linalg.solve(A, b.ravel())
    x, st, it, res = T.solve(b, np.zeros(T.shape3))
    err = np.max(np.abs(x.ravel() - x_ref)) / np.max(np.abs(x_ref))
    if solver == 1:
        assert err < 1e-10, err
    else:
        assert err < 1e-6, err   # defect correction: linear contraction over its 31-sweep budget


def test_standing_wave_dirichlet_to_neumann_known_answer():
    """Linear standing wave: impose the surface potential, solve, compare W with a*omega*cos(kx)."""
    nx, nz, L, h0, a = 65, 13, 2 * math.pi, 1.0, 1e-3
    pr = P()
    bath = pr.flat_bottom(nx, 1, L, 1.0)
    p = oracle_params(nx, 1, nz, lx=L, ly=1.0, nonlinear=0, reltol=1e-13)
    T = oracle.OracleTank(p, bath["h"], bath["hx"], bath["hy"], bath["hxx"], bath["hyy"])
    x, _, _ = T.grid()
    k = 2 * 2 * math.pi / L
    om = math.sqrt(9.81 * k * math.tanh(k * h0))
    Phi = (9.81 * a / om) * np.cos(k * x)[None, :]
    b = np.zeros(T.shape3)
    b[0, 1:-1, -1] = Phi[0, 1:-1]
    phi, st, it, res = T.solve(b, np.zeros(T.shape3))
    W = T.vertical_velocity(phi)
    exact = a * om * np.cos(k * x)
    err = np.max(np.abs(W[0, 1:-1] - exact[1:-1])) / (a * om)
    assert err < 2e-4, err


def test_rk4_linear_standing_wave_one_period():
    nx, nz, L, h0, a = 65, 13, 2 * math.pi, 1.0, 1e-4
    pr = P()
    bath = pr.flat_bottom(nx, 1, L, 1.0)
    p = oracle_params(nx, 1, nz, lx=L, ly=1.0, nonlinear=0, reltol=1e-13)
    T = oracle.OracleTank(p, bath["h"], bath["hx"], bath["hy"], bath["hxx"], bath["hyy"])
    x, _, _ = T.grid()
    k = 2 * 2 * math.pi / L
    om = math.sqrt(9.81 * k * math.tanh(k * h0))
    E0 = (a * np.cos(k * x))[None, :]
    T.set_state(E0, np.zeros_like(E0))
    period = 2 * math.pi / om
    nsteps = 64
    dt = period / nsteps
    for s in range(nsteps):
        T.rk4_step(s * dt, dt)
    E, _, _, _ = T.get_state()
    err = np.max(np.abs(E[0, 1:-1] - E0[0, 1:-1])) / a
    assert err < 5e-3, err
