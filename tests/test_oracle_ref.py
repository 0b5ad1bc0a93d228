"""The C oracle against the reference's own Fortran, compiled from /root/reference by oracle/ref.mk
(oracle/_ref).  Every comparison is exact (the oracle follows the reference's evaluation order and
both are built without FP contraction); tests skip where /root/reference is absent.

Covered: grid and sigma levels, Fornberg stencils (x, y, sigma), the plane ghost update and plane
derivatives (DifferentiationsFreeSurfacePlane), sigma coefficients (DetermineTransformationConstantsArray),
the high-order matvec (BuildLinearSystem incl. wall, corner and bottom rows and the corner-corrected
mixed derivatives), VerticalFreeSurfaceVelocity, rhsFreeSurface3D, the multigrid level matrices
(BuildLinearSystemMatrix after duplicate summing), GaussSeidel, GaussSeidelSpecialPoints,
MGRestriction3DExplicit and MGProlongation3DExplicit.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from conftest import oracle_params

import oracle

pytestmark = pytest.mark.skipif(not (oracle.ref_available() or __import__("os").path.isdir("/root/reference/src")),
                                reason="reference tree absent: oracle/_ref cannot be built")


def problems():
    import ow3d.problems as P
    return P


def tank_pair(nx, ny, nz, alpha, shoal, gridz=1, L=10.0, **kw):
    P = problems()
    bath = (P.gaussian_shoal if shoal else P.flat_bottom)(nx, ny, L, L)
    R = oracle.RefTank(nx, ny, nz, L, L, alpha, gridz=gridz, bath=bath)
    p = oracle_params(nx, ny, nz, alpha=alpha, beta=alpha, gamma=alpha, lx=L, ly=L, gridz=gridz, **kw)
    T = oracle.OracleTank(p, bath["h"], bath["hx"], bath["hy"], bath["hxx"], bath["hyy"])
    return R, T, bath


def surface(nx, ny, L, seed):
    P = problems()
    x, y = P.coords(nx, ny, L, L)
    X, Y = np.meshgrid(x, y)
    rng = np.random.default_rng(seed)
    E = 0.03 * np.sin(2 * np.pi * X / L + 0.4) * (np.cos(2 * np.pi * Y / L) if ny > 1 else 1.0) + 0.003 * rng.standard_normal(X.shape)
    Pp = 0.2 * np.cos(2 * np.pi * X / L) + 0.01 * rng.standard_normal(X.shape)
    return E, Pp


CASES = [(33, 33, 9, 3, True, 1), (33, 33, 9, 3, False, 1), (65, 1, 9, 3, True, 1), (21, 17, 12, 2, True, 1),
         (23, 19, 13, 4, True, 0), (17, 1, 6, 2, False, 0), (9, 9, 5, 2, True, 1)]


@pytest.mark.parametrize("nx,ny,nz,alpha,shoal,gridz", CASES)
def test_highorder_operator_exact(nx, ny, nz, alpha, shoal, gridz):
    R, T, _ = tank_pair(nx, ny, nz, alpha, shoal, gridz=gridz, precond=0)
    for a, b in zip(R.grid(), T.grid()):
        assert np.array_equal(a, b)
    rs, ts = R.stencils(), T.stencils()
    assert np.array_equal(rs[0], ts[0]) and np.array_equal(rs[2], ts[2])
    if ny > 1:
        assert np.array_equal(rs[1], ts[1])
    E, Pp = surface(nx, ny, 10.0, seed=nx + nz)
    Er, Pr, dr = R.surface(E, Pp)
    Eo, Po, do = T.surface(E, Pp)
    assert np.array_equal(Er, Eo) and np.array_equal(Pr, Po)
    for k in ("Ex", "Exx", "Px") + (("Ey", "Eyy", "Py") if ny > 1 else ()):
        assert np.array_equal(dr[k], do[k]), k
    assert np.array_equal(R.dsigma(), T.dsigma())
    phi = np.random.default_rng(3).standard_normal(T.shape3)
    assert np.array_equal(R.matvec(phi), T.matvec(phi))
    W = R.vertical_velocity(phi)
    assert np.array_equal(W, T.vertical_velocity(phi))
    rr = R.rhs(9.81, W)
    ro = T.rhs(Eo, Po, W)
    assert np.array_equal(rr[0], ro[0]) and np.array_equal(rr[1], ro[1])


def _level_bath(T, level):
    """Level coordinates and depths as the oracle builds them (MGPreProcess.f90:216-314)."""
    return T.level_grid(level)


@pytest.mark.parametrize("nx,ny,nz,shoal", [(33, 33, 9, True), (17, 17, 9, True), (65, 1, 9, True), (33, 17, 9, False)])
def test_multigrid_level_matrices(nx, ny, nz, shoal):
    R, T, bath = tank_pair(nx, ny, nz, 3, shoal, mg_mode=0, cd_quirk=1)
    P = problems()
    for level in range(T.nlev):
        lv = T.level_grid(level)
        val, row, col = oracle.ref_lowmatrix(lv["nx"], lv["ny"], lv["nz"], lv["x"], lv["y"], lv["z"], lv)
        n = lv["n"]
        A_ref = sp.coo_matrix((val, (row, col)), shape=(n, n)).tocsr()
        A_ref.sum_duplicates()
        rp, ci, va = T.csr(level)
        A_or = sp.csr_matrix((va, ci, rp), shape=(n, n))
        d = (A_ref - A_or)
        assert abs(d).max() <= 1e-13 * abs(A_ref).max(), (level, abs(d).max())
        if level > 0:
            x = np.random.default_rng(level).standard_normal(n)
            b = np.random.default_rng(level + 7).standard_normal(n)
            # Gauss-Seidel sweeps of the reference on the oracle's CSR (diagonal first) = the oracle's
            xr = oracle.ref_gs(rp, ci, va, x, b)
            xr = oracle.ref_gs(rp, ci, va, xr, b, ighost=T.level_ghosts(level))
            xo = T.smooth_gs_once(level, b, x)
            assert np.max(np.abs(xr - xo)) <= 1e-13 * max(1.0, np.max(np.abs(xr)))


@pytest.mark.parametrize("nx,ny,nz", [(33, 33, 9), (17, 17, 9), (65, 1, 9), (33, 17, 9)])
def test_transfer_operators_exact(nx, ny, nz):
    R, T, _ = tank_pair(nx, ny, nz, 3, True)
    for fine in range(1, T.nlev):
        f = T.level_grid(fine)
        c = T.level_grid(fine - 1)
        rng = np.random.default_rng(fine)
        rf = rng.standard_normal((f["nyg"], f["nxg"], f["nzg"]))
        rc_ref = oracle.ref_transfer(0, (f["nx"], f["ny"], f["nz"]), (f["x"], f["y"], f["z"]),
                                     (c["nx"], c["ny"], c["nz"]), (c["x"], c["y"], c["z"]), rf,
                                     np.zeros((c["nyg"], c["nxg"], c["nzg"])))
        assert np.array_equal(rc_ref, T.restrict(fine, rf))
        ec = rng.standard_normal((c["nyg"], c["nxg"], c["nzg"]))
        ef_ref = oracle.ref_transfer(1, (f["nx"], f["ny"], f["nz"]), (f["x"], f["y"], f["z"]),
                                     (c["nx"], c["ny"], c["nz"]), (c["x"], c["y"], c["z"]),
                                     np.zeros((f["nyg"], f["nxg"], f["nzg"])), ec)
        assert np.array_equal(ef_ref, T.prolong(fine, ec))
