! Test harness (not part of the product): C-callable entry points that drive the reference's
! own Fortran routines for the Laplace path, compiled from /root/reference/src by oracle/ref.mk.
! Used only by tests/ to check the C oracle (oracle/ow3d_oracle.c) against the reference.
MODULE ow3d_ref_driver
  USE ISO_C_BINDING
  USE Precision
  USE Constants
  USE DataTypes
  USE GlobalVariables, ONLY: FineGrid, Lx, Ly, Lz, GridX, GridY, GridZ, GhostGridX, GhostGridY, GhostGridZ, &
       alpha, beta, gamma, Wavefield, swenseONOFF, curvilinearONOFF, g, pi
  IMPLICIT NONE
  INTEGER, PARAMETER :: scratch_unit = 77
  LOGICAL, SAVE :: unit_open = .FALSE.
  TYPE(Level_def), SAVE :: LowGrid
  TYPE(SparseArray_COO), SAVE :: LowMatrix
CONTAINS

  SUBROUTINE open_unit()
    IF (.NOT. unit_open) THEN
       OPEN(unit=scratch_unit, status='scratch')
       unit_open = .TRUE.
    END IF
  END SUBROUTINE open_unit

  ! Grid, stencils and mixed-derivative tables of the high-order operator, as OceanWave3DT0Setup does:
  ! SetupCompDomain, PreProcessDiffStencils, ConstructTableCrossDerivatives(...,0).
  SUBROUTINE ref_setup(nx, ny, nz, lx_, ly_, lz_, gridz_, alp, h, hx, hy, hxx, hyy) BIND(C, name="ref_setup")
    INTEGER(C_INT), VALUE :: nx, ny, nz, gridz_, alp
    REAL(C_DOUBLE), VALUE :: lx_, ly_, lz_
    REAL(C_DOUBLE), DIMENSION(*) :: h, hx, hy, hxx, hyy
    INTEGER :: nxg, nyg, i, j
    CALL open_unit()
    pi = ACOS(-one)
    swenseONOFF = 0
    curvilinearONOFF = 0
    FineGrid%Nx = nx; FineGrid%Ny = ny; FineGrid%Nz = nz
    Lx = lx_; Ly = ly_; Lz = lz_
    GridX = 0; GridY = 0; GridZ = gridz_
    GhostGridX = 1; GhostGridZ = 1
    IF (ny > 1) THEN; GhostGridY = 1; ELSE; GhostGridY = 0; END IF
    alpha = alp; beta = alp; gamma = alp
    IF (ny == 1) beta = 0
    CALL SetupCompDomain
    nxg = nx + 2*GhostGridX; nyg = ny + 2*GhostGridY
    DO j = 1, nyg
       DO i = 1, nxg
          FineGrid%h(i,j) = h(i + (j-1)*nxg)
          FineGrid%hx(i,j) = hx(i + (j-1)*nxg)
          FineGrid%hy(i,j) = hy(i + (j-1)*nxg)
          FineGrid%hxx(i,j) = hxx(i + (j-1)*nxg)
          FineGrid%hyy(i,j) = hyy(i + (j-1)*nxg)
       END DO
    END DO
    CALL PreProcessDiffStencils(FineGrid, FineGrid%DiffStencils, GhostGridX, GhostGridY, GhostGridZ, &
         alpha, beta, gamma, scratch_unit)
    CALL ConstructTableCrossDerivatives(FineGrid, FineGrid%DiffStencils, gamma, GhostGridX, GhostGridY, GhostGridZ, 0)
    CALL ALLOCATE_Wavefield_Type(Wavefield, nx, ny, nz, GhostGridX, GhostGridY, GhostGridZ, 0)
    ALLOCATE(Wavefield%SourceEx(nxg,nyg), Wavefield%SourcePx(nxg,nyg), Wavefield%SourceEy(nxg,nyg), Wavefield%SourcePy(nxg,nyg))
    Wavefield%SourceEx = zero; Wavefield%SourcePx = zero; Wavefield%SourceEy = zero; Wavefield%SourcePy = zero
  END SUBROUTINE ref_setup

  SUBROUTINE ref_grid(x, y, z) BIND(C, name="ref_grid")
    REAL(C_DOUBLE), DIMENSION(*) :: x, y, z
    INTEGER :: i
    DO i = 1, FineGrid%Nx + 2*GhostGridX
       x(i) = FineGrid%x(i, 1)
    END DO
    DO i = 1, FineGrid%Ny + 2*GhostGridY
       y(i) = FineGrid%y(1, i)
    END DO
    DO i = 1, FineGrid%Nz + GhostGridZ
       z(i) = FineGrid%z(i)
    END DO
  END SUBROUTINE ref_grid

  ! Stencil tables in [N][order][r] order (r fastest).
  SUBROUTINE ref_stencils(sx, sy, sz) BIND(C, name="ref_stencils")
    REAL(C_DOUBLE), DIMENSION(*) :: sx, sy, sz
    INTEGER :: i, o, m, r, q
    r = 2*alpha + 1; q = 0
    DO i = 1, FineGrid%Nx + 2*GhostGridX
       DO o = 1, 2
          DO m = 1, r
             q = q + 1; sx(q) = FineGrid%DiffStencils%StencilX(i, m, o)
          END DO
       END DO
    END DO
    IF (FineGrid%Ny > 1) THEN
       r = 2*beta + 1; q = 0
       DO i = 1, FineGrid%Ny + 2*GhostGridY
          DO o = 1, 2
             DO m = 1, r
                q = q + 1; sy(q) = FineGrid%DiffStencils%StencilY(i, m, o)
             END DO
          END DO
       END DO
    END IF
    r = 2*gamma + 1; q = 0
    DO i = 1, FineGrid%Nz + GhostGridZ
       DO o = 1, 2
          DO m = 1, r
             q = q + 1; sz(q) = FineGrid%DiffStencils%StencilZ(i, m, o)
          END DO
       END DO
    END DO
  END SUBROUTINE ref_stencils

  ! DifferentiationsFreeSurfacePlane + DetermineTransformationConstantsArray on (E, P); returns the
  ! ghost-updated E, P and the plane derivatives.
  SUBROUTINE ref_surface(E, P, Ex, Exx, Ey, Eyy, Px, Py) BIND(C, name="ref_surface")
    REAL(C_DOUBLE), DIMENSION(*) :: E, P, Ex, Exx, Ey, Eyy, Px, Py
    INTEGER :: nxg, nyg, i, j, q
    nxg = FineGrid%Nx + 2*GhostGridX; nyg = FineGrid%Ny + 2*GhostGridY
    DO j = 1, nyg
       DO i = 1, nxg
          q = i + (j-1)*nxg
          Wavefield%E(i,j) = E(q); Wavefield%P(i,j) = P(q)
       END DO
    END DO
    CALL DifferentiationsFreeSurfacePlane(Wavefield, GhostGridX, GhostGridY, FineGrid, alpha, beta)
    CALL DetermineTransformationConstantsArray(nxg, nyg, FineGrid%Nz+GhostGridZ, FineGrid, FineGrid%dsigmanew, Wavefield)
    DO j = 1, nyg
       DO i = 1, nxg
          q = i + (j-1)*nxg
          E(q) = Wavefield%E(i,j); P(q) = Wavefield%P(i,j)
          Ex(q) = Wavefield%Ex(i,j); Exx(q) = Wavefield%Exx(i,j); Px(q) = Wavefield%Px(i,j)
          IF (FineGrid%Ny > 1) THEN
             Ey(q) = Wavefield%Ey(i,j); Eyy(q) = Wavefield%Eyy(i,j); Py(q) = Wavefield%Py(i,j)
          END IF
       END DO
    END DO
  END SUBROUTINE ref_surface

  SUBROUTINE ref_dsigma(out) BIND(C, name="ref_dsigma")
    REAL(C_DOUBLE), DIMENSION(*) :: out
    INTEGER :: n, c, i, j, k, q
    n = (FineGrid%Nx+2*GhostGridX)*(FineGrid%Ny+2*GhostGridY)*(FineGrid%Nz+GhostGridZ)
    q = 0
    DO c = 1, 5
       DO j = 1, FineGrid%Ny+2*GhostGridY
          DO i = 1, FineGrid%Nx+2*GhostGridX
             DO k = 1, FineGrid%Nz+GhostGridZ
                q = q + 1
                out(q) = FineGrid%dsigmanew(k,i,j,c)
             END DO
          END DO
       END DO
    END DO
  END SUBROUTINE ref_dsigma

  ! BuildLinearSystem (the high-order matvec)
  SUBROUTINE ref_matvec(phi, out) BIND(C, name="ref_matvec")
    REAL(C_DOUBLE), DIMENSION(*) :: phi, out
    CALL BuildLinearSystem(FineGrid%Nx+2*GhostGridX, FineGrid%Ny+2*GhostGridY, FineGrid%Nz+GhostGridZ, &
         phi, out, FineGrid, alpha, beta, gamma)
  END SUBROUTINE ref_matvec

  SUBROUTINE ref_vertical_velocity(phi, W) BIND(C, name="ref_vertical_velocity")
    REAL(C_DOUBLE), DIMENSION(*) :: phi, W
    CALL VerticalFreeSurfaceVelocity(W, FineGrid%Nx+2*GhostGridX, FineGrid%Ny+2*GhostGridY, FineGrid%Nz+GhostGridZ, &
         phi, FineGrid%DiffStencils, FineGrid%dsigmanew(:,:,:,5), gamma)
  END SUBROUTINE ref_vertical_velocity

  ! rhsFreeSurface3D with the wavefield of the last ref_surface call and the given W
  SUBROUTINE ref_rhs(gr, W, rhsE, rhsP) BIND(C, name="ref_rhs")
    REAL(C_DOUBLE), VALUE :: gr
    REAL(C_DOUBLE), DIMENSION(*) :: W, rhsE, rhsP
    INTEGER :: nxg, nyg, i, j
    nxg = FineGrid%Nx + 2*GhostGridX; nyg = FineGrid%Ny + 2*GhostGridY
    DO j = 1, nyg
       DO i = 1, nxg
          Wavefield%W(i,j) = W(i + (j-1)*nxg)
       END DO
    END DO
    CALL rhsFreeSurface3D(zero, Wavefield, gr, rhsE, rhsP, nxg, nyg)
  END SUBROUTINE ref_rhs

  ! Low-order preconditioner matrix of one multigrid level, before CleanSparseMatrixCOO:
  ! the calls PreparePreconditioner.f90:78-97 makes (Precond = 1, Cartesian), on a level grid
  ! given by its interior counts, coordinates and depths.
  SUBROUTINE ref_lowmatrix(nx, ny, nz, x, y, z, h, hx, hy, hxx, hyy, nnz) BIND(C, name="ref_lowmatrix")
    INTEGER(C_INT), VALUE :: nx, ny, nz
    REAL(C_DOUBLE), DIMENSION(*) :: x, y, z, h, hx, hy, hxx, hyy
    INTEGER(C_INT) :: nnz
    INTEGER :: nxg, nyg, nzg, i, j, a, b, c, gy
    TYPE(Diff_def) :: FullRankStencils
    TYPE(Wavefield_FS) :: tmp_wavefield
    CALL open_unit()
    IF (ASSOCIATED(LowGrid%x)) THEN
       DEALLOCATE(LowGrid%x, LowGrid%y, LowGrid%z, LowGrid%h, LowGrid%hx, LowGrid%hy, LowGrid%hxx, LowGrid%hyy, LowGrid%dsigmanew)
       DEALLOCATE(LowMatrix%val, LowMatrix%row_ptr, LowMatrix%col_ind)
    END IF
    gy = 1; IF (ny == 1) gy = 0
    nxg = nx + 2; nyg = ny + 2*gy; nzg = nz + 1
    LowGrid%Nx = nx; LowGrid%Ny = ny; LowGrid%Nz = nz
    ALLOCATE(LowGrid%x(nxg,nyg), LowGrid%y(nxg,nyg), LowGrid%z(nzg), LowGrid%h(nxg,nyg), LowGrid%hx(nxg,nyg), &
         LowGrid%hy(nxg,nyg), LowGrid%hxx(nxg,nyg), LowGrid%hyy(nxg,nyg), LowGrid%dsigmanew(nzg,nxg,nyg,5))
    LowGrid%dsigmanew = zero
    DO j = 1, nyg
       DO i = 1, nxg
          LowGrid%x(i,j) = x(i); LowGrid%y(i,j) = y(j)
          LowGrid%h(i,j) = h(i + (j-1)*nxg); LowGrid%hx(i,j) = hx(i + (j-1)*nxg)
          LowGrid%hy(i,j) = hy(i + (j-1)*nxg); LowGrid%hxx(i,j) = hxx(i + (j-1)*nxg)
          LowGrid%hyy(i,j) = hyy(i + (j-1)*nxg)
       END DO
    END DO
    DO i = 1, nzg
       LowGrid%z(i) = z(i)
    END DO
    a = 1; b = 1; c = 1
    IF (ny == 1) b = 0
    CALL PreProcessDiffStencils(LowGrid, LowGrid%DiffStencils, 1, gy, 1, a, b, c, scratch_unit)
    CALL PrepareFullOperatorStencils(LowGrid%DiffStencils, FullRankStencils, a, b, c, nxg, nyg, nzg)
    CALL ALLOCATE_Wavefield_Type(tmp_wavefield, nx, ny, nz, 1, gy, 1, 0)
    CALL DetermineTransformationConstantsArray(nxg, nyg, nzg, LowGrid, LowGrid%dsigmanew, tmp_wavefield)
    CALL ConstructTableCrossDerivatives(LowGrid, LowGrid%DiffStencils, c, 1, gy, 1, 1)
    CALL BuildLinearSystemMatrix(nxg, nyg, nzg, LowMatrix, 1, gy, 1, LowGrid%dsigmanew, LowGrid, FullRankStencils, a, b, c)
    nnz = LowMatrix%nnz
  END SUBROUTINE ref_lowmatrix

  SUBROUTINE ref_lowmatrix_get(val, row, col) BIND(C, name="ref_lowmatrix_get")
    REAL(C_DOUBLE), DIMENSION(*) :: val
    INTEGER(C_INT), DIMENSION(*) :: row, col
    INTEGER :: q
    DO q = 1, LowMatrix%nnz
       val(q) = LowMatrix%val(q); row(q) = LowMatrix%row_ptr(q); col(q) = LowMatrix%col_ind(q)
    END DO
  END SUBROUTINE ref_lowmatrix_get

  ! GaussSeidel and GaussSeidelSpecialPoints on a caller-given CSR matrix (1-based indices).
  SUBROUTINE ref_gs(nrow, nnz, rowptr, colind, val, x, b, nghost, ighost) BIND(C, name="ref_gs")
    INTEGER(C_INT), VALUE :: nrow, nnz, nghost
    INTEGER(C_INT), DIMENSION(*) :: rowptr, colind, ighost
    REAL(C_DOUBLE), DIMENSION(*) :: val, x, b
    TYPE(SparseArray_CSR) :: SM
    REAL(KIND=long), ALLOCATABLE :: xx(:), bb(:)
    INTEGER :: q
    SM%nrow = nrow; SM%nnz = nnz
    ALLOCATE(SM%val(nnz), SM%col_ind(nnz), SM%row_ptr(nrow+1), xx(MAX(nnz,nrow)), bb(MAX(nnz,nrow)))
    SM%val = val(1:nnz); SM%col_ind = colind(1:nnz); SM%row_ptr = rowptr(1:nrow+1)
    xx = zero; bb = zero
    DO q = 1, nrow
       xx(q) = x(q); bb(q) = b(q)
    END DO
    IF (nghost > 0) THEN
       CALL GaussSeidelSpecialPoints(xx, bb, SM, ighost(1:nghost), nghost)
    ELSE
       CALL GaussSeidel(xx, bb, SM)
    END IF
    DO q = 1, nrow
       x(q) = xx(q)
    END DO
    DEALLOCATE(SM%val, SM%col_ind, SM%row_ptr, xx, bb)
  END SUBROUTINE ref_gs

  ! MGRestriction3DExplicit / MGProlongation3DExplicit between two level grids given by interior counts
  ! and coordinates.
  SUBROUTINE make_level(L, nx, ny, nz, x, y, z)
    TYPE(Level_def) :: L
    INTEGER :: nx, ny, nz
    REAL(C_DOUBLE), DIMENSION(*) :: x, y, z
    INTEGER :: nxg, nyg, i, j, gy
    gy = 1; IF (ny == 1) gy = 0
    nxg = nx + 2; nyg = ny + 2*gy
    L%Nx = nx; L%Ny = ny; L%Nz = nz
    ALLOCATE(L%x(nxg,nyg), L%y(nxg,nyg), L%z(nz+1))
    DO j = 1, nyg
       DO i = 1, nxg
          L%x(i,j) = x(i); L%y(i,j) = y(j)
       END DO
    END DO
    DO i = 1, nz+1
       L%z(i) = z(i)
    END DO
  END SUBROUTINE make_level

  SUBROUTINE ref_transfer(dir, nxf, nyf, nzf, xf, yf, zf, nxc, nyc, nzc, xc, yc, zc, fine, coarse) BIND(C, name="ref_transfer")
    INTEGER(C_INT), VALUE :: dir, nxf, nyf, nzf, nxc, nyc, nzc
    REAL(C_DOUBLE), DIMENSION(*) :: xf, yf, zf, xc, yc, zc, fine, coarse
    TYPE(Level_def) :: F, C
    INTEGER :: gy
    gy = 1; IF (nyf == 1) gy = 0
    CALL make_level(F, nxf, nyf, nzf, xf, yf, zf)
    CALL make_level(C, nxc, nyc, nzc, xc, yc, zc)
    IF (dir == 0) THEN
       CALL MGRestriction3DExplicit(fine, coarse, F, C, 1, gy, 1, nzf+1, nxf+2, nyf+2*gy, nzc+1, nxc+2, nyc+2*gy)
    ELSE
       CALL MGProlongation3DExplicit(fine, coarse, F, C, 1, gy, 1, nzf+1, nxf+2, nyf+2*gy, nzc+1, nxc+2, nyc+2*gy)
    END IF
    DEALLOCATE(F%x, F%y, F%z, C%x, C%y, C%z)
  END SUBROUTINE ref_transfer


  ! StoreData (IO/StoreData.f90) on the grid of the last ref_setup: writes EP_<nr>.bin in the working
  ! directory, E and P over the ghost-extended plane.
  SUBROUTINE ref_store_ep(E, P, nr) BIND(C, name="ref_store_ep")
    REAL(C_DOUBLE), INTENT(IN) :: E(*), P(*)
    INTEGER(C_INT), VALUE :: nr
    INTEGER :: nxg, nyg
    nxg = FineGrid%Nx + 2*GhostGridX; nyg = FineGrid%Ny + 2*GhostGridY
    CALL StoreData(nxg, nyg, E(1:nxg*nyg), P(1:nxg*nyg), FineGrid, INT(nr), 1)
  END SUBROUTINE ref_store_ep

  ! JONSWAP_spectrum(FREQ, Hs, Tp, gamma) (utilities/JONSWAP_spectrum.f90) at n frequencies in Hz.
  SUBROUTINE ref_jonswap(n, freq, hs, tp, gam, out) BIND(C, name="ref_jonswap")
    INTEGER(C_INT), VALUE :: n
    REAL(C_DOUBLE), VALUE :: hs, tp, gam
    REAL(C_DOUBLE), DIMENSION(*) :: freq, out
    REAL(KIND=long), EXTERNAL :: JONSWAP_spectrum
    REAL(KIND=long) :: f, h, t, gg
    INTEGER :: i
    h = hs; t = tp; gg = gam
    DO i = 1, n
       f = freq(i)
       out(i) = JONSWAP_spectrum(f, h, t, gg)
    END DO
  END SUBROUTINE ref_jonswap

  ! StoreKinematicData(Nx, Ny, Nz, io, it) (IO/StoreKinematicData.f90) with formattype 21 for one output
  ! window: writes Kinematics_01_<it>.bin into the current directory from phi_in and the surface last given
  ! to ref_surface.  win = xbeg, xend, xstride, ybeg, yend, ystride, tbeg, tend, tstride.
  SUBROUTINE ref_store_kinematics(phi_in, win, dt_, nonlin, it) BIND(C, name="ref_store_kinematics")
    USE GlobalVariables, ONLY: PHI, Output, formattype, dt, LinearONOFF
    REAL(C_DOUBLE), DIMENSION(*) :: phi_in
    INTEGER(C_INT), DIMENSION(9) :: win
    REAL(C_DOUBLE), VALUE :: dt_
    INTEGER(C_INT), VALUE :: nonlin, it
    INTEGER :: nxg, nyg, nzg, i, j, k
    nxg = FineGrid%Nx + 2*GhostGridX; nyg = FineGrid%Ny + 2*GhostGridY; nzg = FineGrid%Nz + GhostGridZ
    IF (ASSOCIATED(PHI)) DEALLOCATE(PHI)
    ALLOCATE(PHI(nzg, nxg, nyg))
    DO j = 1, nyg
       DO i = 1, nxg
          DO k = 1, nzg
             PHI(k, i, j) = phi_in(k + nzg*(i-1) + nzg*nxg*(j-1))
          END DO
       END DO
    END DO
    IF (ALLOCATED(Output)) DEALLOCATE(Output)
    ALLOCATE(Output(1))
    Output(1)%xbeg = win(1); Output(1)%xend = win(2); Output(1)%xstride = win(3)
    Output(1)%ybeg = win(4); Output(1)%yend = win(5); Output(1)%ystride = win(6)
    Output(1)%tbeg = win(7); Output(1)%tend = win(8); Output(1)%tstride = win(9)
    formattype = 21; dt = dt_; LinearONOFF = nonlin
    CALL StoreKinematicData(nxg, nyg, nzg, 1, INT(it))
  END SUBROUTINE ref_store_kinematics

END MODULE ow3d_ref_driver
