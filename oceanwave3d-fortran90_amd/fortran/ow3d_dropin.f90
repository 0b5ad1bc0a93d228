! Drop-in replacements with the reference's subroutine names and argument lists, for linking in place
! of functions/BuildLinearSystem.f90, multigrid/MultiGridSolveLinearSystem.f90,
! wrappers/iterative_solution.f90, functions/VerticalFreeSurfaceVelocity.f90,
! timeintegration/DifferentiationsFreeSurfacePlane.f90 and main/rhsFreeSurface3D.f90.  Compiled inside
! the reference tree (they USE its DataTypes / GlobalVariables modules).  The handle is created once by
! ow3d_dropin_init after OceanWave3DT0Setup has built FineGrid; the sigma coefficients of each stage are
! set by DifferentiationsFreeSurfacePlane, which the RK4 driver calls before every solve
! (Runge_Kutta_4.f90:62/117/183/271).  The multigrid runs V cycles only: W and F cycles
! (MultiGridSolveLinearSystem.f90:29-33) stop with a message.
MODULE ow3d_dropin_state
  USE ISO_C_BINDING
  USE ow3d_iface
  IMPLICIT NONE
  TYPE(C_PTR), SAVE :: ow3d_handle = C_NULL_PTR
CONTAINS
  SUBROUTINE ow3d_dropin_init()
    USE GlobalVariables, ONLY: FineGrid, Lx, Ly, Lz, GridZ, alpha, beta, gamma, g, LinearONOFF, solver, Precond, &
         GMRESmaxiterations, reltol, abstol, maxit, extrapolationONOFF, MGmaxgrids, cyclet, PressureTermOnOff, &
         PDampingOnOff
    USE MGLevels, ONLY: nu
    TYPE(ow3d_config) :: c
    INTEGER(C_INT) :: rc
    c%nx = FineGrid%Nx; c%ny = FineGrid%Ny; c%nz = FineGrid%Nz
    c%alpha = alpha; c%beta = MAX(beta, alpha); c%gamma = gamma; c%gridz = GridZ
    c%lx = Lx; c%ly = Ly; c%lz = Lz; c%g = g
    c%nonlinear = LinearONOFF; c%solver = solver; c%precond = Precond; c%gmres_m = GMRESmaxiterations
    c%reltol = reltol; c%abstol = abstol; c%maxit = maxit; c%nu1 = nu(1); c%nu2 = nu(2)
    c%mg_maxgrids = MGmaxgrids; c%extrapolation = extrapolationONOFF; c%mg_precision = 0
    c%pressure_term = PressureTermOnOff; c%pressure_damping = PDampingOnOff
    c%gmres_side = 0   ! left preconditioning, as iterative_solution.f90:20-46
    IF (Precond == 3 .AND. cyclet /= 'V') STOP 'ow3d: only V cycles are implemented (cyclet must be V)'
    rc = ow3d_create(c, FineGrid%h, FineGrid%hx, FineGrid%hy, FineGrid%hxx, FineGrid%hyy, ow3d_handle)
    IF (rc /= 0) STOP 'ow3d_create failed'
  END SUBROUTINE ow3d_dropin_init
END MODULE ow3d_dropin_state

SUBROUTINE BuildLinearSystem(Nx, Ny, Nz, PHI, output, GridStruct, alpha, beta, gamma)
  USE ow3d_dropin_state
  USE DataTypes
  IMPLICIT NONE
  INTEGER :: Nx, Ny, Nz, alpha, beta, gamma
  REAL(C_DOUBLE), DIMENSION(Nx*Ny*Nz), INTENT(IN) :: PHI
  REAL(C_DOUBLE), DIMENSION(Nx*Ny*Nz), INTENT(OUT) :: output
  TYPE(Level_def), INTENT(IN) :: GridStruct
  IF (ow3d_laplace_apply(ow3d_handle, PHI, output) < 0) STOP 'ow3d_laplace_apply failed'
END SUBROUTINE BuildLinearSystem

SUBROUTINE MultiGridSolveLinearSystem(r, u, Nx, Ny, Nz, GhostGridX, GhostGridY, GhostGridZ, cyclet, maxit, reltol, &
     alpha, beta, gamma)
  USE ow3d_dropin_state
  IMPLICIT NONE
  INTEGER :: Nx, Ny, Nz, GhostGridX, GhostGridY, GhostGridZ, maxit, alpha, beta, gamma
  REAL(C_DOUBLE) :: reltol
  CHARACTER(len=1) :: cyclet
  REAL(C_DOUBLE), DIMENSION((Nx+2*GhostGridX)*(Ny+2*GhostGridY)*(Nz+GhostGridZ)), INTENT(IN) :: r
  REAL(C_DOUBLE), DIMENSION((Nx+2*GhostGridX)*(Ny+2*GhostGridY)*(Nz+GhostGridZ)), INTENT(INOUT) :: u
  IF (cyclet /= 'V') STOP 'ow3d: only V cycles are implemented (cyclet must be V)'
  IF (ow3d_precond_apply(ow3d_handle, r, u) < 0) STOP 'ow3d_precond_apply failed'
END SUBROUTINE MultiGridSolveLinearSystem

SUBROUTINE iterative_solution(rhss, n_rows, A_times_x, sol, CurrentGrid)
  USE ow3d_dropin_state
  USE DataTypes
  IMPLICIT NONE
  EXTERNAL A_times_x
  INTEGER :: n_rows
  REAL(C_DOUBLE), DIMENSION(n_rows), INTENT(IN) :: rhss
  REAL(C_DOUBLE), DIMENSION(n_rows), INTENT(INOUT) :: sol
  TYPE(Level_def), INTENT(IN) :: CurrentGrid
  INTEGER(C_INT) :: iters, rc
  REAL(C_DOUBLE) :: res
  rc = ow3d_solve(ow3d_handle, rhss, sol, iters, res)
  IF (rc < 0) STOP 'ow3d_solve failed'
END SUBROUTINE iterative_solution

SUBROUTINE VerticalFreeSurfaceVelocity(W, Nx, Ny, Nz, sol, DiffStencils, dsdz, gamma)
  USE ow3d_dropin_state
  USE DataTypes
  IMPLICIT NONE
  TYPE(Diff_def), INTENT(IN) :: DiffStencils
  INTEGER :: Nx, Ny, Nz, gamma
  REAL(C_DOUBLE), DIMENSION(Nz,Nx,Ny), INTENT(IN) :: dsdz, sol
  REAL(C_DOUBLE), DIMENSION(Nx,Ny), INTENT(OUT) :: W
  IF (ow3d_vertical_velocity(ow3d_handle, sol, W) < 0) STOP 'ow3d_vertical_velocity failed'
END SUBROUTINE VerticalFreeSurfaceVelocity

SUBROUTINE DifferentiationsFreeSurfacePlane(Wavefield, GhostGridX, GhostGridY, FineGrid, alpha, beta)
  USE ow3d_dropin_state
  USE DataTypes
  IMPLICIT NONE
  INTEGER :: GhostGridX, GhostGridY, alpha, beta
  TYPE(Level_def) :: FineGrid
  TYPE(Wavefield_FS) :: Wavefield
  REAL(C_DOUBLE), ALLOCATABLE :: dummy(:, :)
  INTEGER(C_INT) :: rc
  IF (FineGrid%Ny > 1) THEN
     rc = ow3d_set_surface(ow3d_handle, Wavefield%E, Wavefield%P, Wavefield%Ex, Wavefield%Exx, Wavefield%Ey, &
          Wavefield%Eyy, Wavefield%Px, Wavefield%Py)
  ELSE   ! x-z problem: no y derivatives (Wavefield%Ey.. are not allocated)
     ALLOCATE(dummy(SIZE(Wavefield%E, 1), SIZE(Wavefield%E, 2)))
     rc = ow3d_set_surface(ow3d_handle, Wavefield%E, Wavefield%P, Wavefield%Ex, Wavefield%Exx, dummy, dummy, &
          Wavefield%Px, dummy)
     DEALLOCATE(dummy)
  END IF
  IF (rc < 0) STOP 'ow3d_set_surface failed'
END SUBROUTINE DifferentiationsFreeSurfacePlane

SUBROUTINE rhsFreeSurface3D(t, Wavefield, g, rhsE, rhsP, Nx, Ny)
  USE ow3d_dropin_state
  USE DataTypes
  IMPLICIT NONE
  INTEGER :: Nx, Ny
  REAL(C_DOUBLE) :: g, t
  REAL(C_DOUBLE), DIMENSION(Nx, Ny) :: rhsE, rhsP
  TYPE(Wavefield_FS) :: Wavefield
  ! kinematic and dynamic conditions from the plane derivatives of the last DifferentiationsFreeSurfacePlane
  IF (ow3d_rhs_free_surface(ow3d_handle, Wavefield%E, Wavefield%W, rhsE, rhsP) < 0) STOP 'ow3d_rhs_free_surface failed'
  ! applied pressure, breaking and pressure-damping terms (zero unless those features are on)
  IF (Ny == 1) THEN
     rhsE = rhsE - Wavefield%Qr_x
     rhsP = rhsP + Wavefield%P0 - Wavefield%Mr_t
  ELSE
     rhsP = rhsP + Wavefield%P0
  END IF
  rhsE = rhsE + Wavefield%NuD
  rhsP = rhsP + Wavefield%Pd
END SUBROUTINE rhsFreeSurface3D
