! Test program (not part of the product): a Fortran caller of libow3d.so through the ISO_C_BINDING
! interface module ow3d_iface.  It reads one tank's inputs from a stream file written by
! tests/test_fortran_boundary.py, calls ow3d_create, ow3d_set_surface, ow3d_laplace_apply, ow3d_solve and
! ow3d_vertical_velocity, and writes their outputs to a second stream file.
PROGRAM boundary_check
  USE ISO_C_BINDING
  USE ow3d_iface
  IMPLICIT NONE
  TYPE(ow3d_config) :: c
  TYPE(C_PTR) :: tank
  INTEGER(C_INT) :: dims(5), rc, iters, rc_solve
  REAL(C_DOUBLE) :: lens(2), res
  REAL(C_DOUBLE), ALLOCATABLE :: h(:), hx(:), hy(:), hxx(:), hyy(:), E(:), P(:), phi(:), rhs(:)
  REAL(C_DOUBLE), ALLOCATABLE :: Aphi(:), sol(:), W(:), Ex(:), Exx(:), Ey(:), Eyy(:), Px(:), Py(:)
  INTEGER :: nxg, nyg, nzg, np, n
  CHARACTER(len=512) :: fin, fout

  CALL GET_COMMAND_ARGUMENT(1, fin)
  CALL GET_COMMAND_ARGUMENT(2, fout)
  OPEN(10, FILE=TRIM(fin), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
  READ(10) dims, lens
  nxg = dims(1) + 2
  nyg = dims(2) + 2
  IF (dims(2) == 1) nyg = 1
  nzg = dims(3) + 1
  np = nxg * nyg
  n = np * nzg
  ALLOCATE(h(np), hx(np), hy(np), hxx(np), hyy(np), E(np), P(np), phi(n), rhs(n))
  ALLOCATE(Aphi(n), sol(n), W(np), Ex(np), Exx(np), Ey(np), Eyy(np), Px(np), Py(np))
  READ(10) h, hx, hy, hxx, hyy, E, P, phi, rhs
  CLOSE(10)

  c%nx = dims(1); c%ny = dims(2); c%nz = dims(3)
  c%alpha = dims(4); c%beta = dims(4); c%gamma = dims(4); c%gridz = dims(5)
  c%lx = lens(1); c%ly = lens(2); c%lz = 1.0_C_DOUBLE; c%g = 9.81_C_DOUBLE
  c%nonlinear = 1; c%solver = 1; c%precond = 3; c%gmres_m = 23
  c%reltol = 1.0e-10_C_DOUBLE; c%abstol = 0.0_C_DOUBLE
  c%maxit = 1; c%nu1 = 1; c%nu2 = 1; c%mg_maxgrids = 20; c%extrapolation = 1; c%mg_precision = 0
  c%pressure_term = 0; c%pressure_damping = 0; c%gmres_side = 0

  rc = ow3d_create(c, h, hx, hy, hxx, hyy, tank)
  IF (rc /= 0) STOP 'ow3d_create failed'
  rc = ow3d_set_surface(tank, E, P, Ex, Exx, Ey, Eyy, Px, Py)
  IF (rc /= 0) STOP 'ow3d_set_surface failed'
  rc = ow3d_laplace_apply(tank, phi, Aphi)
  IF (rc /= 0) STOP 'ow3d_laplace_apply failed'
  sol = 0.0_C_DOUBLE
  rc_solve = ow3d_solve(tank, rhs, sol, iters, res)
  IF (rc_solve < 0) STOP 'ow3d_solve failed'
  rc = ow3d_vertical_velocity(tank, sol, W)
  IF (rc /= 0) STOP 'ow3d_vertical_velocity failed'
  CALL ow3d_destroy(tank)

  OPEN(11, FILE=TRIM(fout), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(11) rc_solve, iters, res, E, P, Aphi, sol, W
  CLOSE(11)
END PROGRAM boundary_check
