"""Synthetic wave-tank set-ups: bathymetry with analytic derivatives, linear random-wave initial
conditions from a JONSWAP spectrum, and sponge (damping) zones.

These replace the reference's input-deck driven set-up for the benchmark workloads
(bottom files read in SetupInitialConditions.f90:51-111, JONSWAP_spectrum.f90, relaxation zones of
PreprocessRelaxationZones.f90 with GammaFunctions.f90 ftype 9).
"""
import math

import numpy as np

G = 9.81


def coords(nx, ny, lx, ly):
    """Node coordinates incl. one wall ghost layer (SetupCompDomain.f90:35-60)."""
    x = np.array([((i - 1) / (nx - 1)) * lx for i in range(nx + 2)])
    if ny > 1:
        y = np.array([((j - 1) / (ny - 1)) * ly for j in range(ny + 2)])
    else:
        y = np.array([ly])
    return x, y


def wavenumber(omega, h, g=G):
    """Linear dispersion omega^2 = g k tanh(k h), Newton iteration."""
    k = omega * omega / g
    k = max(k, omega / math.sqrt(g * h))
    for _ in range(100):
        t = math.tanh(k * h)
        f = g * k * t - omega * omega
        df = g * t + g * k * h * (1.0 - t * t)
        dk = f / df
        k -= dk
        if abs(dk) < 1e-15 * k:
            break
    return k


def flat_bottom(nx, ny, lx, ly, h0=1.0):
    x, y = coords(nx, ny, lx, ly)
    h = np.full((len(y), len(x)), h0)
    z = np.zeros_like(h)
    return dict(h=h, hx=z.copy(), hy=z.copy(), hxx=z.copy(), hyy=z.copy())


def gaussian_shoal(nx, ny, lx, ly, h0=1.0, amp=0.5, width=0.2):
    """h = h0*(1 - amp*exp(-r^2/(2 width^2))), r = distance from the tank centre / lx."""
    x, y = coords(nx, ny, lx, ly)
    X, Y = np.meshgrid(x, y)
    dx = (X - lx / 2) / lx
    dy = (Y - ly / 2) / lx if ny > 1 else np.zeros_like(X)
    s2 = width * width
    e = np.exp(-(dx * dx + dy * dy) / (2 * s2))
    h = h0 * (1.0 - amp * e)
    # derivatives of -amp*h0*e with respect to x and y (chain rule through dx = (x-lx/2)/lx)
    ex = e * (-dx / s2) / lx
    exx = e * ((dx * dx / s2 - 1.0) / s2) / (lx * lx)
    hx = -amp * h0 * ex
    hxx = -amp * h0 * exx
    if ny > 1:
        ey = e * (-dy / s2) / lx
        eyy = e * ((dy * dy / s2 - 1.0) / s2) / (lx * lx)
        hy = -amp * h0 * ey
        hyy = -amp * h0 * eyy
    else:
        hy = np.zeros_like(h)
        hyy = np.zeros_like(h)
    return dict(h=h, hx=hx, hy=hy, hxx=hxx, hyy=hyy)


def jonswap(omega, omega_p, gamma=3.3):
    """JONSWAP spectral shape (unnormalised, g^2 omega^-5 form)."""
    sig = np.where(omega <= omega_p, 0.07, 0.09)
    r = np.exp(-((omega - omega_p) ** 2) / (2 * sig * sig * omega_p * omega_p))
    return G * G * omega ** -5 * np.exp(-1.25 * (omega_p / omega) ** 4) * gamma ** r


def random_waves(nx, ny, lx, ly, h0=1.0, hs=0.02, kph=1.5, ncomp=64, spread_deg=30.0, gamma=3.3, seed=0):
    """Sum of linear Airy components: eta and surface potential on the grid incl. ghosts."""
    rng = np.random.default_rng(seed)
    kp = kph / h0
    wp = math.sqrt(G * kp * math.tanh(kp * h0))
    om = np.linspace(0.5 * wp, 3.0 * wp, ncomp)
    dw = om[1] - om[0]
    S = jonswap(om, wp, gamma)
    m0 = np.sum(S) * dw
    S *= (hs / 4.0) ** 2 / m0
    amp = np.sqrt(2.0 * S * dw)
    theta = np.deg2rad(rng.uniform(-spread_deg, spread_deg, ncomp)) if ny > 1 else np.zeros(ncomp)
    phase = rng.uniform(0.0, 2 * np.pi, ncomp)
    x, y = coords(nx, ny, lx, ly)
    X, Y = np.meshgrid(x, y)
    E = np.zeros_like(X)
    P = np.zeros_like(X)
    for a, w, th, ph in zip(amp, om, theta, phase):
        k = wavenumber(w, h0)
        arg = k * (np.cos(th) * X + np.sin(th) * Y) + ph
        E += a * np.cos(arg)
        P += (G * a / w) * np.sin(arg)
    return E, P, dict(kp=kp, omega_p=wp, cp=wp / kp, lambda_p=2 * np.pi / kp)


def sponge_gamma(n, p=3.5):
    """ftype 9 profile (GammaFunctions.f90:19-20): 1 - t^p over the zone, t from 0 to 1."""
    t = np.linspace(0.0, 1.0, n) if n > 1 else np.zeros(1)
    return 1.0 - t ** p


def sponge_zones(nx, ny, lx, ly, width):
    """Damping zones of the given physical width along every wall: list of
    (axis, i0, i1, j0, j1, gam) with 0-based inclusive plane indices of the ghost-extended grid;
    profiles reach zero at the wall (reversed for the west/south zones, ftype -9)."""
    x, y = coords(nx, ny, lx, ly)
    zones = []
    ilo = int(np.searchsorted(x, width - 1e-12))       # first node with x >= width
    ihi = int(np.searchsorted(x, lx - width - 1e-12))  # first node with x >= lx - width
    jr = (1, len(y) - 2) if ny > 1 else (0, 0)
    g_w = sponge_gamma(ilo - 1 + 1)[::-1]
    zones.append((0, 1, ilo, jr[0], jr[1], g_w))
    g_e = sponge_gamma(nx - ihi + 1)
    zones.append((0, ihi, nx, jr[0], jr[1], g_e))
    if ny > 1:
        jlo = int(np.searchsorted(y, width - 1e-12))
        jhi = int(np.searchsorted(y, ly - width - 1e-12))
        zones.append((1, 1, nx, 1, jlo, sponge_gamma(jlo)[::-1]))
        zones.append((1, 1, nx, jhi, ny, sponge_gamma(ny - jhi + 1)))
    return zones


def workload(name, seed=0):
    """Tank definitions of the benchmark workloads: returns (cfg, bathymetry dict, E, P, dt, zones)."""
    if name == "tank257":          # 257x257x9, flat bottom, 64-component JONSWAP, 6th order
        nx = ny = 257; nz = 9
        E0, P0, info = None, None, None
        kp = 1.5; lam = 2 * np.pi / kp
        lx = ly = 16 * lam
        bath = flat_bottom(nx, ny, lx, ly)
        E, P, info = random_waves(nx, ny, lx, ly, hs=0.02, kph=1.5, seed=seed)
        dx = lx / (nx - 1)
        dt = 0.5 * dx / info["cp"]
        cfg = dict(nx=nx, ny=ny, nz=nz, lx=lx, ly=ly)
        return cfg, bath, E, P, dt, []
    if name == "shoal513":         # 513x513x9, Gaussian shoal, 5-wavelength sponges, 6th order
        nx = ny = 513; nz = 9
        kp = 1.5; lam = 2 * np.pi / kp
        lx = ly = 32 * lam
        bath = gaussian_shoal(nx, ny, lx, ly)
        E, P, info = random_waves(nx, ny, lx, ly, hs=0.02, kph=1.5, seed=seed)
        dx = lx / (nx - 1)
        dt = 0.5 * dx / info["cp"]
        cfg = dict(nx=nx, ny=ny, nz=nz, lx=lx, ly=ly)
        return cfg, bath, E, P, dt, sponge_zones(nx, ny, lx, ly, 5 * lam)
    if name == "shoal257":         # same family as shoal513 at the same resolution, a quarter of the area
        nx = ny = 257; nz = 9
        kp = 1.5; lam = 2 * np.pi / kp
        lx = ly = 16 * lam
        bath = gaussian_shoal(nx, ny, lx, ly)
        E, P, info = random_waves(nx, ny, lx, ly, hs=0.02, kph=1.5, seed=seed)
        dx = lx / (nx - 1)
        dt = 0.5 * dx / info["cp"]
        cfg = dict(nx=nx, ny=ny, nz=nz, lx=lx, ly=ly)
        return cfg, bath, E, P, dt, sponge_zones(nx, ny, lx, ly, 5 * lam)
    raise KeyError(name)


BATCH_SIZES = (65, 129, 257, 513)
BATCH_NZ = (6, 9, 12)
BATCH_ORDERS = (4, 6, 8)


def batch_draw(seed, bottom=None, sizes=BATCH_SIZES, nzs=BATCH_NZ, orders=BATCH_ORDERS):
    """Scalar parameters of one tank of the many-tank workload, drawn from its own seed: Nx and Ny
    independently from `sizes`, Nz from `nzs`, stencil order from `orders` (redrawn while the sigma
    stencil, order + 1 points, is wider than the Nz + 1 levels: order 8 needs Nz >= 8), depth
    h0 ~ U[0.5, 2], k_p h0 ~ U[0.5, 3], Hs / h0 ~ U[0.005, 0.03]; bottom "flat", "shoal" (Gaussian shoal of
    half the depth) or None (drawn)."""
    rng = np.random.default_rng(seed)
    nx = int(rng.choice(sizes))
    ny = int(rng.choice(sizes))
    nz = int(rng.choice(nzs))
    order = int(rng.choice(orders))
    while order + 1 > nz + 1:
        order = int(rng.choice(orders))
    h0 = float(rng.uniform(0.5, 2.0))
    kph = float(rng.uniform(0.5, 3.0))
    hs = float(rng.uniform(0.005, 0.03)) * h0
    if bottom is None:
        bottom = "shoal" if rng.uniform() < 0.5 else "flat"
    return dict(nx=nx, ny=ny, nz=nz, order=order, h0=h0, kph=kph, hs=hs, bottom=bottom,
                wave_seed=int(rng.integers(1 << 31)))


def batch_cfg(d):
    """Tank configuration of a draw: 16 points per peak wavelength (L = (N - 1) / 16 * lambda_p); GMRES(31) for
    8th-order tanks, whose solves often stop at the 23-step limit unconverged, GMRES(23) otherwise."""
    lam = 2 * np.pi * d["h0"] / d["kph"]
    a = d["order"] // 2
    return dict(nx=d["nx"], ny=d["ny"], nz=d["nz"], lx=(d["nx"] - 1) / 16 * lam, ly=(d["ny"] - 1) / 16 * lam,
                alpha=a, beta=a, gamma=a, gmres_m=31 if d["order"] >= 8 else 23)


def batch_tank_from(d):
    """(cfg, bathymetry, E, P, dt, zones) of a draw: JONSWAP gamma = 3.3, 64 components, dt = 0.5 dx / c_p,
    walls, no sponge."""
    cfg = batch_cfg(d)
    nx, ny, lx, ly = cfg["nx"], cfg["ny"], cfg["lx"], cfg["ly"]
    h0 = d["h0"]
    bath = gaussian_shoal(nx, ny, lx, ly, h0=h0) if d["bottom"] == "shoal" else flat_bottom(nx, ny, lx, ly, h0=h0)
    E, P, info = random_waves(nx, ny, lx, ly, h0=h0, hs=d["hs"], kph=d["kph"], seed=d["wave_seed"])
    dt = 0.5 * (lx / (nx - 1)) / info["cp"]
    return cfg, bath, E, P, dt, []


def batch_tank(seed, bottom=None, sizes=BATCH_SIZES, nzs=BATCH_NZ, orders=BATCH_ORDERS):
    """One tank of the many-tank workload (batch_draw, then batch_tank_from)."""
    return batch_tank_from(batch_draw(seed, bottom, sizes, nzs, orders))
