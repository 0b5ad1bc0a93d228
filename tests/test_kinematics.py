"""Volume kinematics (StoreKinematicData.f90:266-326): u, v, w at every grid point from phi.

The oracle's restatement is pinned by convergence to an exact harmonic potential on a tank with a
non-flat free surface and a shoaling bottom, where the chain-rule terms of the sigma transform carry an
O(1) part of u and v.  A dropped term, a wrong sign, the wrong sigma coordinate or a missing 1/d does not
converge.  The GPU kernel is held to the oracle."""
import numpy as np
import pytest

from conftest import oracle_params


def _exact(nx, ny, nz, order, L=6.0):
    """Tank, surface (level at the walls) and the exact potential phi = cosh(kap (z + 1.3)) cos(k1 x)
    cos(k2 y) on the sigma grid, with its exact velocities (phi_x, phi_y, phi_z at fixed z)."""
    from ow3d import problems
    x, y = problems.coords(nx, ny, L, L)
    bath = problems.gaussian_shoal(nx, ny, L, L)
    X, Y = np.meshgrid(x, y)
    # zero slope at the walls, as the surface's ghost update imposes (UpdateGhostLayerX.f90:15)
    eta = 0.08 * np.cos(2 * np.pi * X / L) * np.cos(2 * np.pi * Y / L)
    k1, k2 = 2 * np.pi / L, 4 * np.pi / L
    kap = np.hypot(k1, k2)
    zs = np.sin(np.pi * np.arange(nz) / (2 * (nz - 1)))       # GridZ = 1
    sig = np.concatenate([[-zs[1]], zs])                        # ghost level: z(1) = -z(3)
    d = bath["h"] + eta
    Z = sig[None, None, :] * d[:, :, None] - bath["h"][:, :, None]
    C, S = np.cos(k1 * X)[:, :, None], np.sin(k1 * X)[:, :, None]
    Cy, Sy = np.cos(k2 * Y)[:, :, None], np.sin(k2 * Y)[:, :, None]
    ch, sh = np.cosh(kap * (Z + 1.3)), np.sinh(kap * (Z + 1.3))
    phi = ch * C * Cy
    u = -k1 * ch * S * Cy
    v = -k2 * ch * C * Sy
    w = kap * sh * C * Cy
    return bath, eta, phi, (u, v, w)


def _oracle_velocities(oracle_mod, nx, ny, nz, order):
    bath, eta, phi, exact = _exact(nx, ny, nz, order)
    a = order // 2
    p = oracle_params(nx, ny, nz, alpha=a, beta=a, gamma=a, lx=6.0, ly=6.0, precond=0)
    T = oracle_mod.OracleTank(p, bath["h"], bath["hx"], bath["hy"], bath["hxx"], bath["hyy"])
    E, P, _ = T.surface(eta, np.zeros_like(eta))
    got = T.kinematics(phi, E)
    return got, exact


def _interior_error(got, exact):
    """Largest error over columns two or more nodes from every wall (the wall ghost values of eta come
    from its Neumann update), levels 1 .. top."""
    err = 0.0
    for g, e in zip(got, exact):
        err = max(err, float(np.max(np.abs(g[2:-2, 2:-2, 1:] - e[2:-2, 2:-2, 1:]))))
    return err


@pytest.mark.parametrize("order", [4, 6])
def test_oracle_kinematics_converge_to_the_exact_velocities(oracle_mod, order):
    errs = []
    for (n, nz) in ((17, 9), (33, 17), (65, 33)):
        got, exact = _oracle_velocities(oracle_mod, n, n, nz, order)
        errs.append(_interior_error(got, exact))
    rates = [np.log2(errs[q] / errs[q + 1]) for q in range(len(errs) - 1)]
    assert errs[-1] < {4: 4e-3, 6: 4e-4}[order], errs
    assert min(rates) > order - 1.3, (errs, rates)


def test_oracle_kinematics_chain_rule_is_essential(oracle_mod):
    """The same field read without the chain-rule terms is wrong at O(eta_x): the test above would catch
    a dropped term."""
    got, exact = _oracle_velocities(oracle_mod, 33, 33, 17, 6)
    assert _interior_error(got, exact) < 1e-2
    bath, eta, phi, _ = _exact(33, 33, 17, 6)
    p = oracle_params(33, 33, 17, alpha=3, beta=3, gamma=3, lx=6.0, ly=6.0, precond=0, nonlinear=0)
    T = oracle_mod.OracleTank(p, bath["h"], bath["hx"], bath["hy"], bath["hxx"], bath["hyy"])
    E, _, _ = T.surface(eta, np.zeros_like(eta))
    lin = T.kinematics(phi, E)          # linear mode: no chain-rule terms (LinearOnOff = 0)
    assert _interior_error(lin[:2], exact[:2]) > 50 * _interior_error(got[:2], exact[:2])
    assert np.array_equal(lin[2], got[2])   # w = W / d either way


@pytest.mark.gpu
@pytest.mark.parametrize("nx,ny,nz,order", [(33, 33, 9, 6), (65, 1, 9, 6), (33, 17, 12, 8)])
def test_gpu_kinematics_match_the_oracle(oracle_mod, nx, ny, nz, order):
    """Same state on both sides (the surface derivatives follow the reference's order: eta_x before the
    y ghost update), then the same two RK4 steps on both sides."""
    import ow3d
    from ow3d import problems
    a = order // 2
    L = 10.0
    bath = problems.gaussian_shoal(nx, ny, L, L)
    x, y = problems.coords(nx, ny, L, L)
    X, Y = np.meshgrid(x, y)
    E = 0.02 * np.sin(2 * np.pi * X / L) * np.cos(2 * np.pi * Y / L)
    P = 0.01 * np.cos(2 * np.pi * X / L)
    p = oracle_params(nx, ny, nz, alpha=a, beta=a, gamma=a, lx=L, ly=L, mg_maxgrids=0, reltol=1e-13)
    T_o = oracle_mod.OracleTank(p, bath["h"], bath["hx"], bath["hy"], bath["hxx"], bath["hyy"])
    cfg = {k: v for k, v in p.items() if k not in ("mg_mode", "cd_quirk")}
    T_g = ow3d.Tank(cfg, **bath)
    phi = np.random.default_rng(4).standard_normal(T_g.shape3) * 0.01
    T_g.set_state(E, P, None, phi)
    T_o.set_state(E, P, None, phi)
    E_o, _, _, _ = T_o.get_state()
    for g, o in zip(T_g.kinematics(), T_o.kinematics(phi, E_o)):
        assert np.max(np.abs(g - o)) <= 1e-12 * max(np.max(np.abs(o)), 1e-300)
    T_g.set_state(E, P)
    T_o.set_state(E, P)
    T_g.rk4_steps(2, 0.0, 0.05)
    for s in range(2):
        T_o.rk4_step(s * 0.05, 0.05)
    Eg, _, _, phig = T_g.get_state()
    Eo, _, _, phio = T_o.get_state()
    for g, o in zip(T_g.kinematics(), T_o.kinematics(phio, Eo)):
        assert np.max(np.abs(g - o)) <= 1e-9 * max(np.max(np.abs(o)), 1e-300)
    # an explicit phi gives the same as the resident one
    assert np.array_equal(T_g.kinematics(phig)[0], T_g.kinematics()[0])


def _exact_gradients(nx, ny, nz, L=6.0):
    """The 9 fields StoreKinematicData.f90:327-478 writes after u, v, w, exact for the potential of _exact:
    d w_z, d w_x + d_x w, d w_y + d_y w (derivatives of W = dphi/dsigma = d w), u_z, u_x, u_y, v_z, v_x, v_y."""
    from ow3d import problems
    bath, eta, phi, (u, v, w) = _exact(nx, ny, nz, 6, L)
    x, y = problems.coords(nx, ny, L, L)
    X, Y = np.meshgrid(x, y)
    k1, k2 = 2 * np.pi / L, 4 * np.pi / L
    kap = np.hypot(k1, k2)
    etax = -0.08 * k1 * np.sin(k1 * X) * np.cos(k1 * Y)
    etay = -0.08 * k1 * np.cos(k1 * X) * np.sin(k1 * Y)
    zs = np.sin(np.pi * np.arange(nz) / (2 * (nz - 1)))
    sig = np.concatenate([[-zs[1]], zs])
    d = (bath["h"] + eta)[:, :, None]
    Z = sig[None, None, :] * d - bath["h"][:, :, None]
    C, S = np.cos(k1 * X)[:, :, None], np.sin(k1 * X)[:, :, None]
    Cy, Sy = np.cos(k2 * Y)[:, :, None], np.sin(k2 * Y)[:, :, None]
    ch, sh = np.cosh(kap * (Z + 1.3)), np.sinh(kap * (Z + 1.3))
    dx, dy = (bath["hx"] + etax)[:, :, None], (bath["hy"] + etay)[:, :, None]
    G = [d * kap ** 2 * ch * C * Cy, d * (-k1 * kap * sh * S * Cy) + dx * w, d * (-k2 * kap * sh * C * Sy) + dy * w,
         -k1 * kap * sh * S * Cy, -k1 ** 2 * ch * C * Cy, k1 * k2 * ch * S * Sy,
         -k2 * kap * sh * C * Sy, k1 * k2 * ch * S * Sy, -k2 ** 2 * ch * C * Cy]
    return bath, eta, phi, G


@pytest.mark.parametrize("order", [4, 6])
def test_oracle_kinematics_gradients_converge_to_the_exact_values(oracle_mod, order):
    """All nine gradient fields converge under refinement at close to the stencil order (second derivatives
    composed of first-derivative passes, so one order below the velocities); a dropped chain-rule term, a
    missing 1/d, or gradients of w instead of W = d w would leave an O(1) error."""
    a = order // 2
    errs = []
    for (n, nz) in ((17, 9), (33, 17), (65, 33)):
        bath, eta, phi, G = _exact_gradients(n, n, nz)
        p = oracle_params(n, n, nz, alpha=a, beta=a, gamma=a, lx=6.0, ly=6.0, precond=0)
        T = oracle_mod.OracleTank(p, bath["h"], bath["hx"], bath["hy"], bath["hxx"], bath["hyy"])
        E, _, _ = T.surface(eta, np.zeros_like(eta))
        got = T.kinematics_gradients(phi, E)
        errs.append([float(np.max(np.abs(got[q][2:-2, 2:-2, 1:] - G[q][2:-2, 2:-2, 1:]))) for q in range(9)])
    errs = np.array(errs)
    rates = np.log2(errs[:-1] / errs[1:])
    assert np.all(errs[-1] < {4: 2e-2, 6: 6e-4}[order]), errs[-1]
    assert np.all(rates[-1] > order - 2.5), rates


@pytest.mark.gpu
@pytest.mark.parametrize("nx,ny,nz,order", [(33, 33, 9, 6), (65, 1, 9, 6), (33, 17, 12, 8)])
def test_gpu_kinematics_gradients_match_the_oracle(oracle_mod, nx, ny, nz, order):
    import ow3d
    from ow3d import problems
    a = order // 2
    L = 10.0
    bath = problems.gaussian_shoal(nx, ny, L, L)
    x, y = problems.coords(nx, ny, L, L)
    X, Y = np.meshgrid(x, y)
    E = 0.02 * np.sin(2 * np.pi * X / L) * np.cos(2 * np.pi * Y / L)
    P = 0.01 * np.cos(2 * np.pi * X / L)
    p = oracle_params(nx, ny, nz, alpha=a, beta=a, gamma=a, lx=L, ly=L, mg_maxgrids=0)
    T_o = oracle_mod.OracleTank(p, bath["h"], bath["hx"], bath["hy"], bath["hxx"], bath["hyy"])
    T_g = ow3d.Tank({k: v for k, v in p.items() if k not in ("mg_mode", "cd_quirk")}, **bath)
    phi = np.random.default_rng(5).standard_normal(T_g.shape3) * 0.01
    T_g.set_state(E, P, None, phi)
    T_o.set_state(E, P, None, phi)
    E_o, _, _, _ = T_o.get_state()
    Gg, Go = T_g.kinematics_gradients(), T_o.kinematics_gradients(phi, E_o)
    for q in range(9):
        assert np.max(np.abs(Gg[q] - Go[q])) <= 1e-12 * max(np.max(np.abs(Go[q])), 1e-300), q
    assert np.array_equal(T_g.kinematics_gradients(phi), Gg)
