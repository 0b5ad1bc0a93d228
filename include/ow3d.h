/*
 * ow3d.h -- C ABI of the MI355X (gfx950) implementation of OceanWave3D's
 * Laplace-solve-per-RK4-stage hot path.
 *
 * Arrays are host arrays owned by the caller unless the function name ends in
 * _device.  Layout is the reference's: 3D fields are sigma-fastest,
 *     phi[k + nzg*(i + nxg*j)],  nxg = Nx+2, nyg = Ny+2 (1 when Ny == 1), nzg = Nz+1,
 * 2D free-surface fields are E[i + nxg*j].  All values are FP64.
 *
 * Return codes: 0 ok, >0 the iterative solver did not reach its tolerance
 * (the reference ignores this, iterative_solution.f90:115-134; here it is reported),
 * <0 error (ow3d_last_error() gives the message).
 *
 * One handle per host thread at a time; each handle owns its device buffers and its HIP stream.
 * Handles may be created, destroyed and have state copied on one thread while others step theirs:
 * graph capture, creation, destruction and state copies are serialised by a process-wide lock (creation
 * releases it while it computes the coarsest multigrid level's dense inverse on the host).
 */
#ifndef OW3D_H
#define OW3D_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct ow3d_tank ow3d_tank;

/* Mirrors the reference's input-deck knobs for the path (ReadInputFileParameters.f90):
 * line 3 (grid), line 4 (half widths), line 7 (solver), plus the precision choice. */
typedef struct ow3d_config {
    int nx, ny, nz;            /* Nx, Ny, Nz (deck line 3); Ny == 1 selects the x-z problem */
    int alpha, beta, gamma;    /* half widths of the high-order stencils (alpha == beta == gamma) */
    int gridz;                 /* 0 even sigma levels, 1 sine-stretched (GridZ) */
    double lx, ly, lz;         /* domain lengths; lz used only for gridz == 0 */
    double g;                  /* gravitational acceleration */
    int nonlinear;             /* 1 fully nonlinear free-surface conditions, 0 linear */
    int solver;                /* 0 defect correction (dc.f), 1 GMRES */
    int precond;               /* 0 none, 3 geometric multigrid */
    int gmres_m;               /* GMRESmaxiterations: Krylov dimension and matvec budget */
    double reltol, abstol;     /* stopping test on the preconditioned residual */
    int maxit;                 /* multigrid cycles per preconditioner application */
    int nu1, nu2;              /* pre / post smoothing sweeps */
    int mg_maxgrids;           /* cap on multigrid levels (<= 0: no cap) */
    int extrapolation;         /* extrapolationONOFF: initial guesses from earlier solutions */
    int mg_precision;          /* 0 FP64 multigrid, 1 FP32 multigrid, 2 FP32 multigrid with the coarsest level's dense inverse stored in FP16 */
    int pressure_term;         /* PressureTermOnOff: applied free-surface pressure of funPressureTerm.f90 (only 0 is supported) */
    int pressure_damping;      /* PDampingOnOff: pressure damping zones of funPressureTerm.f90 (only 0 is supported) */
    int gmres_side;            /* GMRES preconditioning: 0 left (the reference's, iterative_solution.f90:20-46), 1 flexible
                                  right (FGMRES: stops on the FP64 true residual, whose attainable accuracy does not depend
                                  on how exactly an FP32 multigrid is linear; takes more iterations) */
} ow3d_config;

/* Replaces the set-up the reference does in OceanWave3DT0Setup.f90:32-409 for this path
 * (SetupCompDomain, PreProcessDiffStencils, MGPreProcess): grids, stencils and the multigrid
 * hierarchy.  h, hx, hy, hxx, hyy are plane arrays (nxg*nyg); hx.. may be NULL (zero).
 * A configuration asking for the applied pressure or pressure damping terms of funPressureTerm.f90
 * (pressure_term or pressure_damping non-zero; P0, Pd, NuD are zero otherwise) is refused. */
int  ow3d_create(const ow3d_config *cfg, const double *h, const double *hx, const double *hy,
                 const double *hxx, const double *hyy, ow3d_tank **out);
void ow3d_destroy(ow3d_tank *t);
int  ow3d_dims(const ow3d_tank *t, int *nxg, int *nyg, int *nzg, int *mg_levels);
const char *ow3d_last_error(void);

/* Replaces BuildLinearSystem(Nx,Ny,Nz,PHI,output,GridStruct,alpha,beta,gamma)
 * (functions/BuildLinearSystem.f90:1-18): y = A phi with the sigma coefficients of the
 * surface last given to ow3d_set_surface. */
int ow3d_laplace_apply(ow3d_tank *t, const double *phi, double *y);

/* Replaces MultiGridSolveLinearSystem(r,u,Nx,Ny,Nz,...) (multigrid/MultiGridSolveLinearSystem.f90:1-15):
 * u = M^-1 r, maxit multigrid V cycles from u = 0. */
int ow3d_precond_apply(ow3d_tank *t, const double *r, double *u);

/* Replaces iterative_solution(rhss,n_rows,A_times_x,sol,CurrentGrid)
 * (wrappers/iterative_solution.f90:1-16) with A_times_x = BuildLinearSystem: solves A sol = rhs
 * from the initial guess in sol.  iters = matrix-vector products, res = final preconditioned
 * residual norm (estimate for GMRES). */
int ow3d_solve(ow3d_tank *t, const double *rhs, double *sol, int *iters, double *res);

/* ow3d_solve that allocates no device memory (no counterpart in the reference; for callers that cannot
 * allocate: tight request loops, harnesses that forbid hipMalloc).  Same arguments, results and return
 * codes as ow3d_solve, bit for bit.  Its workspace -- reduction partial sums, the staged rhs and solution,
 * the residual and correction vectors and the Krylov basis (and, with flexible GMRES, the preconditioned
 * basis), (GMRESmaxiterations + 5) or (2 GMRESmaxiterations + 5) vectors of the grid -- lives in one static
 * 64 MiB device array per GPU, held until the process ends; a tank whose workspace does not fit (e.g.
 * 257 x 257 x 9 at GMRESmaxiterations 23) is refused with an error.  Every call on a GPU shares that GPU's
 * array: calls that overlap on one GPU (several host threads) race. */
int ow3d_solve_noalloc(ow3d_tank *t, const double *rhs, double *sol, int *iters, double *res);

/* Replaces DifferentiationsFreeSurfacePlane(Wavefield,...) (timeintegration/DifferentiationsFreeSurfacePlane.f90:1-32)
 * followed by DetermineTransformationConstantsArray (functions/DetermineTransformationConstantsArray.f90):
 * updates the wall ghost values of E and P in place and sets the operator's sigma coefficients.
 * Derivative outputs may be NULL. */
int ow3d_set_surface(ow3d_tank *t, double *E, double *P, double *Ex, double *Exx, double *Ey,
                     double *Eyy, double *Px, double *Py);

/* Replaces VerticalFreeSurfaceVelocity(W,Nx,Ny,Nz,sol,DiffStencils,dsdz,gamma)
 * (functions/VerticalFreeSurfaceVelocity.f90:1-20). */
int ow3d_vertical_velocity(ow3d_tank *t, const double *phi, double *W);

/* Replaces rhsFreeSurface3D(t,Wavefield,g,rhsE,rhsP,Nx,Ny) (main/rhsFreeSurface3D.f90:1-29) and
 * rhsLinearFreeSurface3D (main/rhsLinearFreeSurface3D.f90:1-13), using the plane derivatives of the
 * surface last given to ow3d_set_surface. */
int ow3d_rhs_free_surface(ow3d_tank *t, const double *E, const double *W, double *rhsE, double *rhsP);

/* Damping (sponge) zone of RelaxationModule_new.f90 (WavegenONOFF = 0): E, P <- gam*E, gam*P over
 * plane indices [i0,i1] x [j0,j1] (0-based, inclusive); gam runs along x (axis 0) or y (axis 1). */
int ow3d_add_relax_zone(ow3d_tank *t, int axis, int i0, int i1, int j0, int j1, const double *gam);

/* Device-resident time stepping: replaces Runge_Kutta_4(rhsFreeSurface) (timeintegration/Runge_Kutta_4.f90:1-299)
 * driven by OceanWave3DTakeATimeStep.  State is copied in once, stepped on the device, copied out on demand. */
int ow3d_set_state(ow3d_tank *t, const double *E, const double *P, const double *W, const double *phi);
int ow3d_get_state(ow3d_tank *t, double *E, double *P, double *W, double *phi);
int ow3d_rk4_steps(ow3d_tank *t, int nsteps, double time0, double dt);
/* iteration statistics as in iterative_solution.f90:117-134: total, solves, min, max */
int ow3d_stats(const ow3d_tank *t, long long *total, long long *solves, int *itmin, int *itmax);

/* Volume kinematics: replaces the velocity part of StoreKinematicData (IO/StoreKinematicData.f90:266-326).
 * U, V = horizontal velocities, W = vertical velocity at every grid point (3D layout) from phi (NULL: the
 * handle's current solution) and the handle's current surface. */
int ow3d_kinematics(ow3d_tank *t, const double *phi, double *U, double *V, double *W);
/* Replaces the velocity-gradient part of StoreKinematicData (IO/StoreKinematicData.f90:327-478): G receives
 * 9 fields of nxg*nyg*nzg values in the order the reference writes them -- Dsigma(W)/d, (W)_x, (W)_y for
 * W = dphi/dsigma (that is d w_z, (d w)_x, (d w)_y), then u_z, u_x, u_y, v_z, v_x, v_y -- from phi (NULL: the
 * handle's current solution) and the handle's current surface.  Uses the Krylov basis as workspace
 * (GMRESmaxiterations >= 12). */
int ow3d_kinematics_gradients(ow3d_tank *t, const double *phi, double *G);
/* The surface slopes eta_x, eta_y (nxg*nyg each) StoreKinematicData writes after eta
 * (IO/StoreKinematicData.f90:244-255): those of the handle's state, computed as Runge_Kutta_4.f90:394
 * computes them at the end of each step (and by ow3d_set_state), with the ghost values the surface had
 * before its own ghost update; ow3d_set_surface replaces them with the slopes of the surface it is given. */
int ow3d_surface_slopes(ow3d_tank *t, double *Ex, double *Ey);

/* Device-pointer form of ow3d_laplace_apply for callers that keep phi resident (stream: hipStream_t
 * or NULL for the handle's stream).  d_phi must be 16-byte aligned (as hipMalloc returns), d_y 8-byte
 * aligned; otherwise the call fails with an error.  The operator's sigma coefficients are rewritten on the
 * handle's stream by surface updates and RK4 stages: on another stream the matvec is ordered after the
 * handle's work enqueued so far, and the handle's later work after the matvec.  Fails while asynchronous
 * steps are in flight. */
int ow3d_laplace_apply_device(ow3d_tank *t, const double *d_phi, double *d_y, void *stream);
/* The handle's HIP stream. */
void *ow3d_stream(ow3d_tank *t);
/* Average duration (ms) of the high-order matvec kernel over the last n_launches launches timed with
 * HIP events on the handle's stream (ow3d_time_matvec runs them). */
int ow3d_time_matvec(ow3d_tank *t, int n_launches, double *avg_ms);

/* Many tanks at once (no counterpart in the reference, which advances one tank per process): every
 * handle's solves run on the device with no host round trip, and from a handle's second step on each
 * RK4 step is one replay of a captured hipGraph on the handle's stream, so tanks advance concurrently.
 * ow3d_batch_rk4_steps advances tanks[i] by nsteps steps of dt[i], each tank on its own stream as soon
 * as its previous step has settled, and returns when all are done (1 if any solve missed its tolerance). */
int ow3d_batch_rk4_steps(ow3d_tank *const *tanks, int ntanks, int nsteps, const double *dt);
/* Asynchronous form for request loops: ow3d_rk4_steps_async enqueues the first of nsteps steps and
 * returns; each ow3d_query call settles completed steps and enqueues the next, returning 1 when all
 * nsteps are done, 0 while steps remain; ow3d_wait blocks and returns as ow3d_rk4_steps. */
int ow3d_rk4_steps_async(ow3d_tank *t, int nsteps, double dt);
int ow3d_query(ow3d_tank *t);
int ow3d_wait(ow3d_tank *t);
/* 1 if a Laplace solve of the steps enqueued by the last ow3d_rk4_steps / ow3d_rk4_steps_async /
 * ow3d_batch_rk4_steps call missed its tolerance, 0 if all converged; an error while steps are in flight.
 * (The reference ignores non-convergence, iterative_solution.f90:115-134.) */
int ow3d_steps_status(ow3d_tank *t);
/* Each stage's solve is enqueued with the iteration slots recent steps used plus a margin; a step in
 * which a solve used all its slots without stopping is rolled back and replayed with the whole budget,
 * so results equal those of whole-budget steps.  Stages are captured per (stage, slots).
 * Replays, graph captures and the host time spent capturing, so far: */
int ow3d_step_stats(const ow3d_tank *t, long long *replays, long long *captures, double *capture_ms);
/* Launch choices of a handle, set by the autotuner (ow3d/autotune.py; no counterpart in the reference):
 * matvec_variant 0 column kernel with the widest tile fitting 64 KiB of LDS, 1 column kernel with 16 x 8
 * tiles, 2 point kernel, 3 column kernel with each column split over two waves; sweep_threads 0
 * automatic (1024 below 256 tiles, else 256), 256 or 1024.
 * Results do not depend on them. */
int ow3d_tune(ow3d_tank *t, int matvec_variant, int sweep_threads);
/* Device memory held by a handle, and what ow3d_create would allocate for cfg (admission control). */
long long ow3d_device_bytes(const ow3d_tank *t);
long long ow3d_estimate_bytes(const ow3d_config *cfg);
/* Multigrid hierarchy ow3d_create would build for cfg: number of levels and unknowns of the coarsest
 * (solved with its dense inverse; at most 4096).  Host only. */
int ow3d_mg_plan(const ow3d_config *cfg, int *nlev, int *coarsest);

#ifdef __cplusplus
}
#endif
#endif
