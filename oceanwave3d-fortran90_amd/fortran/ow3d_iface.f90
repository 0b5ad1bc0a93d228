! ISO_C_BINDING interface to libow3d.so (include/ow3d.h) for Fortran callers.
! Arrays are passed by reference in the reference's own layout: PHI(Nz+1, Nx+2, Ny+2) sigma-fastest,
! plane fields E(Nx+2, Ny+2).  Every function returns 0 on success, > 0 when the iterative solver
! stopped at its budget, < 0 on error.
MODULE ow3d_iface
  USE ISO_C_BINDING
  IMPLICIT NONE

  TYPE, BIND(C) :: ow3d_config
     INTEGER(C_INT)  :: nx, ny, nz, alpha, beta, gamma, gridz
     REAL(C_DOUBLE)  :: lx, ly, lz, g
     INTEGER(C_INT)  :: nonlinear, solver, precond, gmres_m
     REAL(C_DOUBLE)  :: reltol, abstol
     INTEGER(C_INT)  :: maxit, nu1, nu2, mg_maxgrids, extrapolation, mg_precision
     INTEGER(C_INT)  :: pressure_term, pressure_damping, gmres_side
  END TYPE ow3d_config

  INTERFACE
     INTEGER(C_INT) FUNCTION ow3d_create(cfg, h, hx, hy, hxx, hyy, tank) BIND(C, name="ow3d_create")
       IMPORT :: C_INT, C_DOUBLE, C_PTR, ow3d_config
       TYPE(ow3d_config), INTENT(IN) :: cfg
       REAL(C_DOUBLE), DIMENSION(*), INTENT(IN) :: h, hx, hy, hxx, hyy
       TYPE(C_PTR), INTENT(OUT) :: tank
     END FUNCTION ow3d_create
     SUBROUTINE ow3d_destroy(tank) BIND(C, name="ow3d_destroy")
       IMPORT :: C_PTR
       TYPE(C_PTR), VALUE :: tank
     END SUBROUTINE ow3d_destroy
     INTEGER(C_INT) FUNCTION ow3d_laplace_apply(tank, phi, y) BIND(C, name="ow3d_laplace_apply")
       IMPORT :: C_INT, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: tank
       REAL(C_DOUBLE), DIMENSION(*), INTENT(IN) :: phi
       REAL(C_DOUBLE), DIMENSION(*), INTENT(OUT) :: y
     END FUNCTION ow3d_laplace_apply
     INTEGER(C_INT) FUNCTION ow3d_precond_apply(tank, r, u) BIND(C, name="ow3d_precond_apply")
       IMPORT :: C_INT, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: tank
       REAL(C_DOUBLE), DIMENSION(*), INTENT(IN) :: r
       REAL(C_DOUBLE), DIMENSION(*), INTENT(OUT) :: u
     END FUNCTION ow3d_precond_apply
     INTEGER(C_INT) FUNCTION ow3d_solve(tank, rhs, sol, iters, res) BIND(C, name="ow3d_solve")
       IMPORT :: C_INT, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: tank
       REAL(C_DOUBLE), DIMENSION(*), INTENT(IN) :: rhs
       REAL(C_DOUBLE), DIMENSION(*), INTENT(INOUT) :: sol
       INTEGER(C_INT), INTENT(OUT) :: iters
       REAL(C_DOUBLE), INTENT(OUT) :: res
     END FUNCTION ow3d_solve
     INTEGER(C_INT) FUNCTION ow3d_set_surface(tank, E, P, Ex, Exx, Ey, Eyy, Px, Py) BIND(C, name="ow3d_set_surface")
       IMPORT :: C_INT, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: tank
       REAL(C_DOUBLE), DIMENSION(*), INTENT(INOUT) :: E, P
       REAL(C_DOUBLE), DIMENSION(*), INTENT(OUT) :: Ex, Exx, Ey, Eyy, Px, Py
     END FUNCTION ow3d_set_surface
     INTEGER(C_INT) FUNCTION ow3d_vertical_velocity(tank, phi, W) BIND(C, name="ow3d_vertical_velocity")
       IMPORT :: C_INT, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: tank
       REAL(C_DOUBLE), DIMENSION(*), INTENT(IN) :: phi
       REAL(C_DOUBLE), DIMENSION(*), INTENT(OUT) :: W
     END FUNCTION ow3d_vertical_velocity
     INTEGER(C_INT) FUNCTION ow3d_rhs_free_surface(tank, E, W, rhsE, rhsP) BIND(C, name="ow3d_rhs_free_surface")
       IMPORT :: C_INT, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: tank
       REAL(C_DOUBLE), DIMENSION(*), INTENT(IN) :: E, W
       REAL(C_DOUBLE), DIMENSION(*), INTENT(OUT) :: rhsE, rhsP
     END FUNCTION ow3d_rhs_free_surface
     INTEGER(C_INT) FUNCTION ow3d_add_relax_zone(tank, axis, i0, i1, j0, j1, gam) BIND(C, name="ow3d_add_relax_zone")
       IMPORT :: C_INT, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: tank
       INTEGER(C_INT), VALUE :: axis, i0, i1, j0, j1
       REAL(C_DOUBLE), DIMENSION(*), INTENT(IN) :: gam
     END FUNCTION ow3d_add_relax_zone
     INTEGER(C_INT) FUNCTION ow3d_set_state(tank, E, P, W, phi) BIND(C, name="ow3d_set_state")
       IMPORT :: C_INT, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: tank
       REAL(C_DOUBLE), DIMENSION(*), INTENT(IN) :: E, P, W, phi
     END FUNCTION ow3d_set_state
     INTEGER(C_INT) FUNCTION ow3d_get_state(tank, E, P, W, phi) BIND(C, name="ow3d_get_state")
       IMPORT :: C_INT, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: tank
       REAL(C_DOUBLE), DIMENSION(*), INTENT(OUT) :: E, P, W, phi
     END FUNCTION ow3d_get_state
     INTEGER(C_INT) FUNCTION ow3d_rk4_steps(tank, nsteps, time0, dt) BIND(C, name="ow3d_rk4_steps")
       IMPORT :: C_INT, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: tank
       INTEGER(C_INT), VALUE :: nsteps
       REAL(C_DOUBLE), VALUE :: time0, dt
     END FUNCTION ow3d_rk4_steps
  END INTERFACE
END MODULE ow3d_iface
