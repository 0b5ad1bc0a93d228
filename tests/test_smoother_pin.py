"""Pin of the oracle's zebra-line smoother (the GPU multigrid is held to it) by brute force.

The GPU smooths with z-line red-black Gauss-Seidel instead of the reference's lexicographic point
Gauss-Seidel (multigrid/GaussSeidel.f90:15-26).  A red (or black) pass must equal, column by column,
the exact solve of that sigma-column's diagonal block of the level matrix with every other column
frozen, where the level matrix is the reference-ordered CSR assembly without the cross-derivative
replication (BuildLinearSystemMatrix.f90:37-324, cd_quirk = 0).  The only couplings left out are the
documented same-colour ones: the corner-shifted x (y) window of the mixed derivative on the sigma-ghost
level (ConstructTableCrossDerivatives.f90:173-186) reaches the column two away, which has the same
colour.  The wall-ghost pass must leave every ghost row's residual at zero and touch nothing else.

The level operators are the linearised (eta = 0) ones built at set-up, so the surface plays no part.
A dropped or mis-signed coupling term in the smoother changes a block solve at O(1) and fails here.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from conftest import oracle_params

import oracle


def _tank(nx, ny, nz, order, shoal=True):
    import ow3d.problems as pr
    L = 10.0
    A = order // 2
    bath = (pr.gaussian_shoal if shoal else pr.flat_bottom)(nx, ny, L, L)
    p = oracle_params(nx, ny, nz, alpha=A, beta=A, gamma=A, lx=L, ly=L, mg_mode=0, cd_quirk=0)
    return oracle.OracleTank(p, bath["h"], bath["hx"], bath["hy"], bath["hxx"], bath["hyy"])


def _matrix(T, level):
    rp, ci, va = T.csr(level)
    n = len(rp) - 1
    return sp.csr_matrix((va, ci, rp), shape=(n, n))


def _expected_colour(A, x, b, dims, colour, gx, gy):
    """Block Gauss-Seidel over the columns of one colour, built from the assembled matrix."""
    nxg, nyg, nzg = dims
    n = A.shape[0]
    col = np.arange(n) // nzg
    ci, cj = col % nxg, col // nxg
    interior = (ci >= gx) & (ci < nxg - gx) & (cj >= gy) & (cj < nyg - gy)
    parity = (ci + cj) & 1
    xn = x.copy()
    dropped = []
    for c in range(nxg * nyg):
        i, j = c % nxg, c // nxg
        if not (gx <= i < nxg - gx and gy <= j < nyg - gy) or ((i + j) & 1) != colour:
            continue
        rows = np.arange(c * nzg, (c + 1) * nzg)
        blk = A[rows].tocoo()
        own = col[blk.col] == c
        same = interior[blk.col] & (parity[blk.col] == colour) & ~own
        for q in np.nonzero(same)[0]:
            dropped.append((ci[blk.col[q]] - i, cj[blk.col[q]] - j))
        Acc = np.zeros((nzg, nzg))
        np.add.at(Acc, (blk.row[own], blk.col[own] - c * nzg), blk.data[own])
        keep = ~own & ~same
        rhs = b[rows].copy()
        np.add.at(rhs, blk.row[keep], -blk.data[keep] * x[blk.col[keep]])
        xn[rows] = np.linalg.solve(Acc, rhs)
    return xn, dropped


CASES = [(9, 9, 5, 4), (17, 17, 9, 6), (17, 1, 9, 6), (13, 11, 8, 6)]


@pytest.mark.parametrize("nx,ny,nz,order", CASES)
def test_zebra_colours_are_exact_column_block_solves(nx, ny, nz, order):
    T = _tank(nx, ny, nz, order)
    gx, gy = 1, (1 if ny > 1 else 0)
    rng = np.random.default_rng(nx * 100 + nz)
    total_dropped = 0
    for level in range(1, T.nlev):
        d = T.level_dims(level)
        A = _matrix(T, level)
        x0 = rng.standard_normal(A.shape[0])
        b = rng.standard_normal(A.shape[0])
        shape = (d[1], d[0], d[2])
        x = x0
        for colour in (0, 1):
            want, dropped = _expected_colour(A, x, b, d, colour, gx, gy)
            got = T.zebra(level, b.reshape(shape), x.reshape(shape), colour).ravel()
            err = np.max(np.abs(got - want)) / np.max(np.abs(want))
            assert err < 1e-11, (level, colour, err)
            # the only couplings a colour pass leaves out: same colour, two columns away along x or y
            assert all(abs(di) + abs(dj) == 2 and di * dj == 0 for di, dj in dropped), set(dropped)
            total_dropped += len(dropped)
            x = want
        # wall-ghost pass: ghost rows solved, nothing else touched
        got = T.zebra(level, b.reshape(shape), x.reshape(shape), 2).ravel()
        nxg, nyg, nzg = d
        colidx = np.arange(A.shape[0]) // nzg
        i, j = colidx % nxg, colidx // nxg
        wall = (i == 0) | (i == nxg - 1) | ((j == 0) | (j == nyg - 1) if gy else False)
        assert np.array_equal(got[~wall], x[~wall])
        r = (b - A @ got)[wall]
        assert np.max(np.abs(r)) < 1e-11 * max(1.0, np.max(np.abs(b))), np.max(np.abs(r))
    assert total_dropped > 0   # the corner-shifted window is present on these levels


@pytest.mark.parametrize("shoal", [False, True])
@pytest.mark.parametrize("mode", [0, 1])
def test_vcycle_contracts_on_flat_and_shoal(mode, shoal):
    """V(1,1) contraction of both smoothers on the level operator, on a flat bottom and on a shoal
    (where the sigma-x / sigma-y cross terms are non-zero)."""
    import ow3d.problems as pr
    nx = ny = 33
    bath = (pr.gaussian_shoal if shoal else pr.flat_bottom)(nx, ny, 10.0, 10.0)
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
