# Reference checker: the reference's own Fortran routines for the Laplace path, compiled from the
# reference tree as they lie (no copies, no stand-ins), plus the harness oracle/ref_driver.f90.
# Output only into oracle/_ref/.  Routines that call libraries absent from the image (SPARSKIT2,
# HSL MA41, LAPACK: iterative_solution, dc, MultiGridSolver, CleanSparseMatrixCOO, ...) are not built.
REFOBJ := _ref/obj
REF_MODS := variabledefs/precision.f90 variabledefs/constants.f90 variabledefs/datatypes.f90 \
            variabledefs/HSL_LU.f90 variabledefs/OFmodule.f90 variabledefs/globalvariables.f90 \
            multigrid/MGLevels.f90
REF_SRCS := functions/BuildLinearSystem.f90 functions/DiffXEven.f90 functions/DiffYEven.f90 \
            functions/DiffZArbitrary.f90 functions/DiffXEven_CD.f90 functions/DiffYEven_CD.f90 \
            functions/DiffZArbitrary_CD.f90 functions/ConstructTableCrossDerivatives.f90 \
            functions/DiffStencilsZ_modif.f90 functions/DetermineTransformationConstantsArray.f90 \
            functions/VerticalFreeSurfaceVelocity.f90 functions/UpdateGhostLayerX.f90 \
            functions/UpdateGhostLayerY.f90 functions/ALLOCATE_Wavefield_Type.f90 \
            functions/PrepareFullOperatorStencils.f90 functions/BuildLinearSystemMatrix.f90 \
            functions/TaylorFDStencils1DArbitrary.f90 initialization/SetupCompDomain.f90 \
            initialization/PreProcessDiffStencils.f90 initialization/BuildStencilsGridX.f90 \
            initialization/BuildStencilsGridY.f90 utilities/stream_func_wave_finite.f \
            timeintegration/DifferentiationsFreeSurfacePlane.f90 main/rhsFreeSurface3D.f90 \
            multigrid/GaussSeidel.f90 multigrid/GaussSeidelSpecialPoints.f90 \
            multigrid/MGRestriction3DExplicit.f90 multigrid/MGProlongation3DExplicit.f90 IO/CheckError.f90 \
            IO/StoreData.f90 utilities/JONSWAP_spectrum.f90 IO/StoreKinematicData.f90

ref: _ref/libow3d_ref.so

_ref/libow3d_ref.so: ref_driver.f90 ref.mk
	@mkdir -p $(REFOBJ)
	@set -e; for f in $(REF_MODS); do \
	  $(FC) $(FFLAGS) -c $(REFSRC)/$$f -module-dir $(REFOBJ) -o $(REFOBJ)/$$(basename $$f).o; done
	@set -e; for f in $(REF_SRCS); do \
	  $(FC) $(FFLAGS) -c $(REFSRC)/$$f -module-dir $(REFOBJ) -I$(REFOBJ) -o $(REFOBJ)/$$(basename $$f).o; done
	$(FC) $(FFLAGS) -c ref_driver.f90 -module-dir $(REFOBJ) -I$(REFOBJ) -o $(REFOBJ)/ref_driver.o
	$(FC) -shared -o $@ $(REFOBJ)/*.o
