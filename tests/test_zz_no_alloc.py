"""ow3d_solve_noalloc (workspace in the library's static device array) against ow3d_solve on the same
inputs: the same solution bits, iteration count, residual estimate and status, for GMRES and defect
correction, on two calls in a row with the second on a smaller tank (the array then still holds the first
call's contents); a tank whose workspace exceeds the array is refused with an error."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _tank(n, nz, solver, prec=0):
    import ow3d
    from ow3d import problems
    L = 10.0
    bath = problems.gaussian_shoal(n, n, L, L)
    T = ow3d.Tank(dict(nx=n, ny=n, nz=nz, lx=L, ly=L, solver=solver, reltol=1e-10, gmres_m=23, mg_precision=prec,
                       gmres_side=prec), **bath)
    x, y = problems.coords(n, n, L, L)
    X, Y = np.meshgrid(x, y)
    E = 0.02 * np.sin(2 * np.pi * X / L + 0.3) * np.cos(2 * np.pi * Y / L)
    P = 0.3 * np.cos(2 * np.pi * X / L) * np.sin(2 * np.pi * Y / L + 0.2)
    E, P, _ = T.set_surface(E, P)
    b = np.zeros(T.shape3)
    b[1:-1, 1:-1, -1] = P[1:-1, 1:-1]
    x0 = 0.01 * np.random.default_rng(n).standard_normal(T.shape3)
    return T, b, x0


@pytest.mark.parametrize("solver,prec", [(1, 0), (0, 0), (1, 1)])
def test_noalloc_solve_equals_solve(solver, prec):
    """GMRES (left preconditioning with the FP64 multigrid; flexible with the FP32 one, whose workspace
    also holds the preconditioned basis) and defect correction."""
    for n, nz in ((65, 9), (33, 12)):   # second call on a smaller tank
        print(f"no-alloc solve: {n}x{n}x{nz}, solver {solver}, multigrid precision {prec}", flush=True)
        T, b, x0 = _tank(n, nz, solver, prec)
        got = T.solve_noalloc(b, x0)
        want = T.solve(b, x0)
        assert np.array_equal(got[0], want[0]), (n, nz)
        assert got[1:] == want[1:], (got[1:], want[1:])
        assert np.all(np.isfinite(got[0])) and np.max(np.abs(got[0])) > 0
        T.close()


def test_noalloc_solve_refuses_a_tank_beyond_the_array():
    import ow3d
    T, b, x0 = _tank(257, 9, 1)
    with pytest.raises(ow3d.OW3DError, match="too large for the no-allocation solve"):
        T.solve_noalloc(b, x0)
    T.close()
