/*
 * ow3d_oracle.c -- CPU restatement of the OceanWave3D Laplace-solve-per-RK4-stage path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path and the timed CPU
 * baseline.  Nothing in the product library links or calls this file.
 *
 * Every function cites the reference routine it restates (paths relative to the
 * reference's src/ directory).  Indices are 0-based here; the reference is 1-based.
 *
 * Parity status of each part (how it is pinned) is recorded in the test files
 * tests/test_oracle_*.py; summary:
 *   - grid, stencils, sigma coefficients, high-order matvec, plane derivatives,
 *     ghost-layer update, vertical velocity, FSBC right-hand side, low-order CSR
 *     assembly, Gauss-Seidel, restriction, prolongation: compared with the reference
 *     Fortran compiled from /root/reference (oracle/_ref) and with closed forms.
 *   - GMRES (SPARSKIT2 is not in the reference tree): pinned by convergence to the
 *     dense direct solution on assembled small systems and by invariants.
 *   - RK4 driver: pinned by the linear test equation and by the linear standing-wave
 *     known answer (TestCodeDucrozet2.f90:364-455 restated as a test).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "ow3d_oracle.h"

#define MAXR 17

/* ------------------------------------------------------------------------- */
/* Grid                                                                       */
/* ------------------------------------------------------------------------- */
typedef struct {
    int nx, ny, nz;       /* interior counts (deck convention) */
    int gx, gy;           /* ghost layers in x / y (0 or 1); the sigma ghost layer is always present */
    int nxg, nyg, nzg, nxy, n;
    int a, b, c;          /* half widths; a = 0 when nx == 1, b = 0 when ny == 1 */
    int ra, rb, rc;       /* stencil ranks 2a+1, 2b+1, 2c+1 */
    double *x, *y, *z;    /* 1D coordinates incl. ghosts */
    double *h, *hx, *hy, *hxx, *hyy;   /* plane fields [nxy] */
    double *sx, *sy, *sz; /* stencil tables [N][2][r] (order 1, order 2) */
} grid_t;

static void *xcalloc(size_t n, size_t s) {
    void *p = calloc(n ? n : 1, s);
    if (!p) { fprintf(stderr, "ow3d_oracle: out of memory\n"); abort(); }
    return p;
}

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* Fornberg's recursion for finite-difference weights on arbitrary nodes
 * (B. Fornberg, Math. Comp. 51 (1988) 699-706), in the form the reference calls
 * `weights` (utilities/stream_func_wave_finite.f:128-168).  c[j*(m+1)+k] is the weight of
 * node j for the k-th derivative at z, j = 0..n, k = 0..m. */
void owo_fornberg(double z, const double *x, int n, int m, double *c) {
#define C(j, k) c[(j) * (m + 1) + (k)]
    for (int j = 0; j <= n; j++)
        for (int k = 0; k <= m; k++) C(j, k) = 0.0;
    double c1 = 1.0, c4 = x[0] - z;
    C(0, 0) = 1.0;
    for (int i = 1; i <= n; i++) {
        int mn = i < m ? i : m;
        double c2 = 1.0, c5 = c4;
        c4 = x[i] - z;
        for (int j = 0; j <= i - 1; j++) {
            double c3 = x[i] - x[j];
            c2 = c2 * c3;
            if (j == i - 1) {
                for (int k = mn; k >= 1; k--)
                    C(i, k) = c1 * ((double)k * C(i - 1, k - 1) - c5 * C(i - 1, k)) / c2;
                C(i, 0) = -(c1 * c5 * C(i - 1, 0) / c2);
            }
            for (int k = mn; k >= 1; k--)
                C(j, k) = (c4 * C(j, k) - (double)k * C(j, k - 1)) / c3;
            C(j, 0) = c4 * C(j, 0) / c3;
        }
        c1 = c2;
    }
#undef C
}

/* Per-node stencil rows for first and second derivatives; the window of node i is
 * [clamp(i-half, 0, N-r), +r) (BuildStencilsGridX.f90:121-150, BuildStencilsGridY.f90,
 * PreProcessDiffStencils.f90:73-86 for sigma). */
static void build_stencils(const double *x, int N, int half, double *S) {
    int r = 2 * half + 1;
    double c[MAXR * 3];
    for (int i = 0; i < N; i++) {
        int ws = clampi(i - half, 0, N - r);
        for (int o = 1; o <= 2; o++) {
            owo_fornberg(x[i], x + ws, r - 1, o, c);
            for (int m = 0; m < r; m++) S[(i * 2 + (o - 1)) * r + m] = c[m * (o + 1) + o];
        }
    }
}

static void grid_alloc(grid_t *g) {
    g->nxg = g->nx + 2 * g->gx;
    g->nyg = g->ny + 2 * g->gy;
    g->nzg = g->nz + 1;
    g->nxy = g->nxg * g->nyg;
    g->n = g->nxy * g->nzg;
    g->ra = 2 * g->a + 1; g->rb = 2 * g->b + 1; g->rc = 2 * g->c + 1;
    g->x = xcalloc(g->nxg, sizeof(double));
    g->y = xcalloc(g->nyg, sizeof(double));
    g->z = xcalloc(g->nzg, sizeof(double));
    g->h = xcalloc(g->nxy, sizeof(double));
    g->hx = xcalloc(g->nxy, sizeof(double));
    g->hy = xcalloc(g->nxy, sizeof(double));
    g->hxx = xcalloc(g->nxy, sizeof(double));
    g->hyy = xcalloc(g->nxy, sizeof(double));
    g->sx = xcalloc((size_t)g->nxg * 2 * g->ra, sizeof(double));
    g->sy = xcalloc((size_t)g->nyg * 2 * g->rb, sizeof(double));
    g->sz = xcalloc((size_t)g->nzg * 2 * g->rc, sizeof(double));
}

static void grid_free(grid_t *g) {
    free(g->x); free(g->y); free(g->z); free(g->h); free(g->hx); free(g->hy);
    free(g->hxx); free(g->hyy); free(g->sx); free(g->sy); free(g->sz);
}

static void grid_stencils(grid_t *g) {
    if (g->nx > 1) build_stencils(g->x, g->nxg, g->a, g->sx);
    if (g->ny > 1) build_stencils(g->y, g->nyg, g->b, g->sy);
    build_stencils(g->z, g->nzg, g->c, g->sz);
}

/* SetupCompDomain.f90:35-79: uniform x/y incl. ghost nodes; sigma levels even or
 * sine-stretched, plus the reflected ghost level z(1) = -z(3). */
static void grid_fine_coords(grid_t *g, double lx, double ly, double lz, int gridz) {
    if (g->nx == 1) g->x[0] = lx;
    else for (int ii = 0; ii < g->nxg; ii++) g->x[ii] = ((double)(ii - g->gx) / (double)(g->nx - 1)) * lx;
    if (g->ny == 1) g->y[0] = ly;
    else for (int jj = 0; jj < g->nyg; jj++) g->y[jj] = ((double)(jj - g->gy) / (double)(g->ny - 1)) * ly;
    for (int k = 0; k < g->nz; k++) {
        if (gridz == 1) g->z[k + 1] = sin(M_PI * (double)k / (double)(2 * (g->nz - 1)));
        else g->z[k + 1] = (double)k / (double)(g->nz - 1) * lz;
    }
    g->z[0] = -g->z[2];
}

#define I3(g, k, i, j) ((k) + (size_t)(g)->nzg * ((i) + (size_t)(g)->nxg * (j)))
#define I2(g, i, j) ((i) + (size_t)(g)->nxg * (j))

/* ------------------------------------------------------------------------- */
/* 1D derivative passes                                                      */
/* ------------------------------------------------------------------------- */
/* DiffXEven.f90:1-40 on a field of depth nz (nz = nzg for 3D fields, 1 for the plane). */
static void diff_x(const grid_t *g, int nz, const double *v, double *out, int order) {
    int r = g->ra;
    for (int j = 0; j < g->nyg; j++)
        for (int i = 0; i < g->nxg; i++) {
            int ws = clampi(i - g->a, 0, g->nxg - r);
            const double *S = g->sx + (i * 2 + order - 1) * r;
            for (int k = 0; k < nz; k++) {
                double s = 0.0;
                for (int m = 0; m < r; m++) s += S[m] * v[k + (size_t)nz * ((ws + m) + (size_t)g->nxg * j)];
                out[k + (size_t)nz * (i + (size_t)g->nxg * j)] = s;
            }
        }
}

/* DiffYEven.f90:41-78 */
static void diff_y(const grid_t *g, int nz, const double *v, double *out, int order) {
    int r = g->rb;
    for (int j = 0; j < g->nyg; j++) {
        int ws = clampi(j - g->b, 0, g->nyg - r);
        const double *S = g->sy + (j * 2 + order - 1) * r;
        for (int i = 0; i < g->nxg; i++)
            for (int k = 0; k < nz; k++) {
                double s = 0.0;
                for (int m = 0; m < r; m++) s += S[m] * v[k + (size_t)nz * (i + (size_t)g->nxg * (ws + m))];
                out[k + (size_t)nz * (i + (size_t)g->nxg * j)] = s;
            }
    }
}

/* DiffZArbitrary.f90:79-114 */
static void diff_z(const grid_t *g, const double *v, double *out, int order) {
    int r = g->rc;
    for (size_t col = 0; col < (size_t)g->nxy; col++) {
        const double *vc = v + col * g->nzg;
        double *oc = out + col * g->nzg;
        for (int k = 0; k < g->nzg; k++) {
            int ws = clampi(k - g->c, 0, g->nzg - r);
            const double *S = g->sz + (k * 2 + order - 1) * r;
            double s = 0.0;
            for (int m = 0; m < r; m++) s += S[m] * vc[ws + m];
            oc[k] = s;
        }
    }
}

/* First pass of the corner-corrected mixed derivative (DiffXEven_CD with the index and
 * stencil tables of ConstructTableCrossDerivatives.f90:100-186,231-235): on the sigma ghost
 * level (k = 0) of a non-wall column, an x-window that touches a wall ghost column is shifted
 * one node inwards and takes the stencil row of the neighbouring node, so that the unsolved
 * bottom-corner ghost points are never read. */
static void diff_x_cd(const grid_t *g, const double *v, double *out) {
    int r = g->ra, N = g->nxg;
    for (int j = 0; j < g->nyg; j++)
        for (int i = 0; i < N; i++) {
            int ws = clampi(i - g->a, 0, N - r);
            for (int k = 0; k < g->nzg; k++) {
                int sh = 0;
                if (k == 0 && i != 0 && i != N - 1) {
                    if (ws == 0) sh = 1;
                    if (ws + r - 1 == N - 1) sh = -1;
                }
                const double *S = g->sx + ((i - sh) * 2 + 0) * r;
                double s = 0.0;
                for (int m = 0; m < r; m++) s += S[m] * v[I3(g, k, ws + sh + m, j)];
                out[I3(g, k, i, j)] = s;
            }
        }
}

/* y counterpart (DiffYEven_CD; ConstructTableCrossDerivatives.f90:90-98,180-185,248-252). */
static void diff_y_cd(const grid_t *g, const double *v, double *out) {
    int r = g->rb, N = g->nyg;
    for (int j = 0; j < N; j++) {
        int ws = clampi(j - g->b, 0, N - r);
        for (int i = 0; i < g->nxg; i++)
            for (int k = 0; k < g->nzg; k++) {
                int sh = 0;
                if (k == 0 && j != 0 && j != N - 1) {
                    if (ws == 0) sh = 1;
                    if (ws + r - 1 == N - 1) sh = -1;
                }
                const double *S = g->sy + ((j - sh) * 2 + 0) * r;
                double s = 0.0;
                for (int m = 0; m < r; m++) s += S[m] * v[I3(g, k, i, ws + sh + m)];
                out[I3(g, k, i, j)] = s;
            }
    }
}

/* ------------------------------------------------------------------------- */
/* Sigma-transformation coefficients                                         */
/* ------------------------------------------------------------------------- */
/* DetermineTransformationConstantsArray.f90:143-205, one point.  zs = sigma (0 on the ghost
 * level, line 143).  Expression shapes follow the reference's evaluation order. */
static inline void sigma3(double zs, double d, double hx, double hy, double hxx, double hyy,
                          double Ex, double Ey, double Exx, double Eyy, double *s) {
    double s1 = zs;
    s[0] = s1;
    s[1] = ((1.0 - s1) * hx - s1 * Ex) / d;
    s[2] = -(2.0 * (1.0 - s1) * (hx * hx + hy * hy) / (d * d))
           + 2.0 * (2.0 * s1 - 1.0) * (Ex * hx + Ey * hy) / (d * d)
           + (1.0 - s1) * (hxx + hyy) / d
           + 2.0 * s1 * (Ex * Ex + Ey * Ey) / (d * d)
           - s1 * (Exx + Eyy) / d;
    s[3] = ((1.0 - s1) * hy - s1 * Ey) / d;
    s[4] = 1.0 / d;
}

static inline void sigma2d(double zs, double d, double hx, double hxx, double Ex, double Exx, double *s) {
    double s1 = zs;
    s[0] = s1;
    s[1] = ((1.0 - s1) * hx - s1 * Ex) / d;
    s[2] = -(2.0 * (1.0 - s1) * (hx * hx) / (d * d))
           + 2.0 * (2.0 * s1 - 1.0) * (Ex * hx) / (d * d)
           + (1.0 - s1) * (hxx) / d
           + 2.0 * s1 * (Ex * Ex) / (d * d)
           - s1 * (Exx) / d;
    s[3] = 0.0;
    s[4] = 1.0 / d;
}

/* Fill ds[c*n + idx] (c = 0..4 <-> dsigma(:,:,:,1..5)). E* may be NULL for the linearised
 * (eta = 0) operator used by the preconditioner (PreparePreconditioner.f90:90-92). */
static void compute_dsigma(const grid_t *g, const double *E, const double *Ex, const double *Exx,
                           const double *Ey, const double *Eyy, double *ds) {
    size_t n = g->n;
    for (int j = 0; j < g->nyg; j++)
        for (int i = 0; i < g->nxg; i++) {
            size_t p = I2(g, i, j);
            double e = E ? E[p] : 0.0;
            double ex = (Ex && g->nx > 1) ? Ex[p] : 0.0, exx = (Exx && g->nx > 1) ? Exx[p] : 0.0;
            double ey = (Ey && g->ny > 1) ? Ey[p] : 0.0, eyy = (Eyy && g->ny > 1) ? Eyy[p] : 0.0;
            double d = e + g->h[p];
            for (int k = 0; k < g->nzg; k++) {
                double zs = (k == 0) ? 0.0 : g->z[k];
                double s[5];
                if (g->ny == 1) sigma2d(zs, d, g->hx[p], g->hxx[p], ex, exx, s);
                else sigma3(zs, d, g->hx[p], g->hy[p], g->hxx[p], g->hyy[p], ex, ey, exx, eyy, s);
                size_t q = I3(g, k, i, j);
                for (int c = 0; c < 5; c++) ds[c * n + q] = s[c];
            }
        }
}

/* ------------------------------------------------------------------------- */
/* Tank                                                                       */
/* ------------------------------------------------------------------------- */
typedef struct {
    grid_t g;
    int xc, yc, zc;          /* this level was coarsened from the next finer one in x / y / z */
    double *ds;              /* linearised sigma coefficients, 5*n */
    /* reference mode: CSR with the diagonal first in each row (CSRdiaREORDER.f90) */
    int *rp, *ci; double *va; int nnz;
    int *ighost; int nghost; /* special points (MGPreProcess.f90:376-439) */
    /* coarsest level: dense LU with partial pivoting of the level matrix */
    double *lu; int *piv;
    double *x, *b, *r, *t1; /* work vectors */
} level_t;

typedef struct { int axis, i0, i1, j0, j1; double *gam; } zone_t;

struct owo_tank {
    owo_params p;
    grid_t g;
    double *ds;                              /* dsigma, 5*n */
    double *E, *P, *W, *Ex, *Exx, *Ey, *Eyy, *Px, *Py;   /* plane state */
    double *phi, *rhs, *last[2];
    double *tmp[10];                         /* matvec temporaries */
    int nlev;
    level_t *lev;                            /* lev[0] coarsest .. lev[nlev-1] finest */
    zone_t *zones; int nzones;
    long tstep;
    long it_total, it_solves, it_min, it_max;
    /* solver workspace */
    double *kw;                              /* Krylov basis (m+1)*n */
    double *kz;                              /* flexible GMRES: preconditioned basis vectors m*n */
    double *v1, *v2, *v3;
};

/* ------------------------------------------------------------------------- */
/* High-order matvec  (BuildLinearSystem.f90:1-290)                           */
/* ------------------------------------------------------------------------- */
void owo_matvec(owo_tank *t, const double *phi, double *out) {
    const grid_t *g = &t->g;
    const double *ds = t->ds;
    size_t n = g->n;
    double *dpdx = t->tmp[0], *ddpdxx = t->tmp[1], *dpdy = t->tmp[2], *ddpdyy = t->tmp[3];
    double *dpds = t->tmp[4], *ddpdss = t->tmp[5], *ddpdsdx = t->tmp[6], *ddpdsdy = t->tmp[7];
    double *tmp = t->tmp[8];
    const int nxg = g->nxg, nyg = g->nyg, nzg = g->nzg;
    memset(out, 0, n * sizeof(double));
    if (g->nx > 1) {
        diff_x(g, nzg, phi, dpdx, 1);
        diff_x(g, nzg, phi, ddpdxx, 2);
        diff_x_cd(g, phi, tmp);
        /* DiffZArbitrary_CD at non-wall columns (the only place the result is read):
         * the sigma window is never shifted there (ConstructTableCrossDerivatives.f90:154-162). */
        diff_z(g, tmp, ddpdsdx, 1);
    }
    if (g->ny > 1) {
        diff_y(g, nzg, phi, dpdy, 1);
        diff_y(g, nzg, phi, ddpdyy, 2);
        diff_y_cd(g, phi, tmp);
        diff_z(g, tmp, ddpdsdy, 1);
    }
    diff_z(g, phi, dpds, 1);
    diff_z(g, phi, ddpdss, 2);
#define DS(c, q) ds[(size_t)(c) * n + (q)]
    if (g->ny == 1) {
        int j = 0;
        for (int i = g->gx; i < nxg - g->gx; i++)
            for (int k = 0; k < nzg; k++) {
                size_t q = I3(g, k, i, j);
                size_t p = I2(g, i, j);
                if (k == nzg - 1) out[q] = phi[q];
                else if (k == 0) {
                    out[q] = g->hx[p] * dpdx[q + 1] + (DS(4, q + 1) + g->hx[p] * DS(1, q + 1)) * dpds[q + 1];
                } else {
                    out[q] = ddpdxx[q] + 2.0 * DS(1, q) * ddpdsdx[q]
                           + (DS(1, q) * DS(1, q) + DS(4, q) * DS(4, q)) * ddpdss[q]
                           + DS(2, q) * dpds[q];
                }
            }
    } else {
        for (int j = g->gy; j < nyg - g->gy; j++)
            for (int i = g->gx; i < nxg - g->gx; i++)
                for (int k = 0; k < nzg; k++) {
                    size_t q = I3(g, k, i, j);
                    size_t p = I2(g, i, j);
                    if (k == nzg - 1) out[q] = phi[q];
                    else if (k == 0) {
                        out[q] = (DS(4, q + 1) + g->hx[p] * DS(1, q + 1) + g->hy[p] * DS(3, q + 1)) * dpds[q + 1]
                               + g->hx[p] * dpdx[q + 1] + g->hy[p] * dpdy[q + 1];
                    } else {
                        out[q] = ddpdxx[q] + ddpdyy[q] + DS(2, q) * dpds[q]
                               + 2.0 * (DS(1, q) * ddpdsdx[q] + DS(3, q) * ddpdsdy[q])
                               + (DS(1, q) * DS(1, q) + DS(3, q) * DS(3, q) + DS(4, q) * DS(4, q)) * ddpdss[q];
                    }
                }
    }
#undef DS
    /* wall ghost columns: Neumann rows d/dx at the adjacent boundary node (lines 150-216) */
    if (g->gx == 1)
        for (int j = g->gy; j < nyg - g->gy; j++)
            for (int k = 0; k < nzg; k++) {
                out[I3(g, k, 0, j)] = dpdx[I3(g, k, 1, j)];
                out[I3(g, k, nxg - 1, j)] = dpdx[I3(g, k, nxg - 2, j)];
            }
    if (g->gy == 1)
        for (int i = g->gx; i < nxg - g->gx; i++)
            for (int k = 0; k < nzg; k++) {
                out[I3(g, k, i, 0)] = dpdy[I3(g, k, i, 1)];
                out[I3(g, k, i, nyg - 1)] = dpdy[I3(g, k, i, nyg - 2)];
            }
    /* identity rows: vertical corner columns and the bottom ghost of wall columns (lines 217-284) */
    if (g->gx + g->gy == 2) {
        int ci[4] = {0, nxg - 1, 0, nxg - 1}, cj[4] = {0, 0, nyg - 1, nyg - 1};
        for (int c = 0; c < 4; c++)
            for (int k = 0; k < nzg; k++) out[I3(g, k, ci[c], cj[c])] = phi[I3(g, k, ci[c], cj[c])];
        for (int i = 0; i < nxg; i++) {
            out[I3(g, 0, i, 0)] = phi[I3(g, 0, i, 0)];
            out[I3(g, 0, i, nyg - 1)] = phi[I3(g, 0, i, nyg - 1)];
        }
        for (int j = 0; j < nyg; j++) {
            out[I3(g, 0, 0, j)] = phi[I3(g, 0, 0, j)];
            out[I3(g, 0, nxg - 1, j)] = phi[I3(g, 0, nxg - 1, j)];
        }
    }
    if (g->gx == 1 && g->ny == 1) {
        out[I3(g, 0, 0, 0)] = phi[I3(g, 0, 0, 0)];
        out[I3(g, 0, nxg - 1, 0)] = phi[I3(g, 0, nxg - 1, 0)];
    }
}

/* ------------------------------------------------------------------------- */
/* Free-surface plane operations                                             */
/* ------------------------------------------------------------------------- */
/* UpdateGhostLayerX.f90:1-22: ghost value such that the boundary first-derivative stencil
 * reproduces the source S (zero for walls). */
static void ghost_x(const grid_t *g, double *v, const double *S) {
    int r = g->ra, N = g->nxg;
    for (int j = g->gy; j < g->nyg - g->gy; j++) {
        const double *st = g->sx + (1 * 2 + 0) * r;
        double s = 0.0;
        for (int m = 1; m < r; m++) s += st[m] * v[I2(g, m, j)];
        v[I2(g, 0, j)] = ((S ? S[I2(g, 1, j)] : 0.0) - s) / st[0];
        st = g->sx + ((N - 2) * 2 + 0) * r;
        s = 0.0;
        for (int m = 0; m < r - 1; m++) s += st[m] * v[I2(g, N - r + m, j)];
        v[I2(g, N - 1, j)] = ((S ? S[I2(g, N - 2, j)] : 0.0) - s) / st[r - 1];
    }
}

/* UpdateGhostLayerY.f90 */
static void ghost_y(const grid_t *g, double *v, const double *S) {
    int r = g->rb, N = g->nyg;
    for (int i = g->gx; i < g->nxg - g->gx; i++) {
        const double *st = g->sy + (1 * 2 + 0) * r;
        double s = 0.0;
        for (int m = 1; m < r; m++) s += st[m] * v[I2(g, i, m)];
        v[I2(g, i, 0)] = ((S ? S[I2(g, i, 1)] : 0.0) - s) / st[0];
        st = g->sy + ((N - 2) * 2 + 0) * r;
        s = 0.0;
        for (int m = 0; m < r - 1; m++) s += st[m] * v[I2(g, i, N - r + m)];
        v[I2(g, i, N - 1)] = ((S ? S[I2(g, i, N - 2)] : 0.0) - s) / st[r - 1];
    }
}

/* DifferentiationsFreeSurfacePlane.f90:12-32 (Cartesian branch), then the sigma coefficients
 * (DetermineTransformationConstantsArray) from the new surface. */
static void surface_update(owo_tank *t, double *E, double *P) {
    const grid_t *g = &t->g;
    if (g->nx > 1) {
        ghost_x(g, E, NULL);
        diff_x(g, 1, E, t->Ex, 1);
        diff_x(g, 1, E, t->Exx, 2);
        ghost_x(g, P, NULL);
        diff_x(g, 1, P, t->Px, 1);
    }
    if (g->ny > 1) {
        ghost_y(g, E, NULL);
        diff_y(g, 1, E, t->Ey, 1);
        diff_y(g, 1, E, t->Eyy, 2);
        ghost_y(g, P, NULL);
        diff_y(g, 1, P, t->Py, 1);
    }
    if (t->p.nonlinear)
        compute_dsigma(g, E, t->Ex, t->Exx, t->Ey, t->Eyy, t->ds);
}

void owo_surface(owo_tank *t, double *E, double *P, double *Ex, double *Exx, double *Ey, double *Eyy,
                 double *Px, double *Py) {
    surface_update(t, E, P);
    size_t b = (size_t)t->g.nxy * sizeof(double);
    if (Ex) memcpy(Ex, t->Ex, b);
    if (Exx) memcpy(Exx, t->Exx, b);
    if (Ey) memcpy(Ey, t->Ey, b);
    if (Eyy) memcpy(Eyy, t->Eyy, b);
    if (Px) memcpy(Px, t->Px, b);
    if (Py) memcpy(Py, t->Py, b);
}

void owo_get_dsigma(const owo_tank *t, double *out) { memcpy(out, t->ds, 5 * (size_t)t->g.n * sizeof(double)); }

/* VerticalFreeSurfaceVelocity.f90:1-20: one-sided sigma derivative at the surface times dsigma/dz. */
void owo_vertical_velocity(owo_tank *t, const double *phi, double *W) {
    const grid_t *g = &t->g;
    int r = g->rc, nzg = g->nzg;
    const double *st = g->sz + ((nzg - 1) * 2 + 0) * r;
    for (size_t col = 0; col < (size_t)g->nxy; col++) {
        double s = 0.0;
        for (int m = 0; m < r; m++) s += st[m] * phi[col * nzg + (nzg - r + m)];
        W[col] = t->ds[4 * (size_t)g->n + col * nzg + (nzg - 1)] * s;
    }
}

/* Fluid velocities at every grid point (ghost points included) from phi, as StoreKinematicData.f90:266-326
 * computes them: W = dphi/dsigma (DiffZArbitrary), U = dphi/dx (DiffXEven) and V = dphi/dy (DiffYEven), in
 * nonlinear mode plus the chain-rule terms ((1 - z)/d h_x - z/d eta_x) W and ((1 - z)/d h_y - z/d eta_y) W
 * with z the level's sigma coordinate (the ghost level's included, StoreKinematicData.f90:290-297,
 * 309-316) and d = h + eta; the vertical velocity is W / d (:326).  eta_x, eta_y are those of the last
 * owo_surface / owo_set_state call, E is that surface. */
void owo_kinematics(owo_tank *t, const double *phi, const double *E, double *U, double *V, double *Wv) {
    const grid_t *g = &t->g;
    const size_t n = g->n;
    double *W = xcalloc(n, sizeof(double));
    diff_z(g, phi, W, 1);
    if (g->nx > 1) diff_x(g, g->nzg, phi, U, 1); else memset(U, 0, n * sizeof(double));
    if (g->ny > 1) diff_y(g, g->nzg, phi, V, 1); else memset(V, 0, n * sizeof(double));
    for (int j = 0; j < g->nyg; j++)
        for (int i = 0; i < g->nxg; i++) {
            const size_t p = I2(g, i, j);
            const double d = g->h[p] + E[p];
            for (int k = 0; k < g->nzg; k++) {
                const size_t q = I3(g, k, i, j);
                const double z = g->z[k];
                if (t->p.nonlinear) {
                    if (g->nx > 1) U[q] = U[q] + ((1 - z) / d * g->hx[p] - z / d * t->Ex[p]) * W[q];
                    if (g->ny > 1) V[q] = V[q] + ((1 - z) / d * g->hy[p] - z / d * t->Ey[p]) * W[q];
                }
                Wv[q] = W[q] / d;
            }
        }
    free(W);
}

/* The velocity gradients StoreKinematicData.f90:327-478 writes after u, v, w, as 9 fields of n in its order:
 *   Dsigma(W)/d, Dx(W) + cx Dsigma(W), Dy(W) + cy Dsigma(W)      with W = dphi/dsigma (= d w)
 *   Dsigma(U)/d, Dx(U) + cx Dsigma(U), Dy(U) + cy Dsigma(U)      (u_z, u_x, u_y)
 *   Dsigma(V)/d, Dx(V) + cx Dsigma(V), Dy(V) + cy Dsigma(V)      (v_z, v_x, v_y)
 * with U, V the velocities of owo_kinematics, cx = (1 - z)/d h_x - z/d eta_x and cy likewise (nonlinear mode
 * only, :337-344, 353-360, ...), x (y) derivatives zero when Nx (Ny) is 1 (:346-348).  The first three are
 * derivatives of d w, not of w: d w_z, (d w)_x, (d w)_y. */
void owo_kinematics_gradients(owo_tank *t, const double *phi, const double *E, double *G) {
    const grid_t *g = &t->g;
    const size_t n = g->n;
    double *U = xcalloc(n, sizeof(double)), *V = xcalloc(n, sizeof(double)), *Wv = xcalloc(n, sizeof(double));
    double *W = xcalloc(n, sizeof(double)), *Fz = xcalloc(n, sizeof(double));
    owo_kinematics(t, phi, E, U, V, Wv);
    diff_z(g, phi, W, 1);
    const double *F[3] = {W, U, V};
    for (int f = 0; f < 3; f++) {
        double *Gz = G + (size_t)(3 * f) * n, *Gx = Gz + n, *Gy = Gx + n;
        diff_z(g, F[f], Fz, 1);
        if (g->nx > 1) diff_x(g, g->nzg, F[f], Gx, 1); else memset(Gx, 0, n * sizeof(double));
        if (g->ny > 1) diff_y(g, g->nzg, F[f], Gy, 1); else memset(Gy, 0, n * sizeof(double));
        for (int j = 0; j < g->nyg; j++)
            for (int i = 0; i < g->nxg; i++) {
                const size_t p = I2(g, i, j);
                const double d = g->h[p] + E[p];
                for (int k = 0; k < g->nzg; k++) {
                    const size_t q = I3(g, k, i, j);
                    const double z = g->z[k];
                    Gz[q] = Fz[q] / d;
                    if (t->p.nonlinear) {
                        if (g->nx > 1) Gx[q] = Gx[q] + ((1 - z) / d * g->hx[p] - z / d * t->Ex[p]) * Fz[q];
                        if (g->ny > 1) Gy[q] = Gy[q] + ((1 - z) / d * g->hy[p] - z / d * t->Ey[p]) * Fz[q];
                    }
                }
            }
    }
    free(U); free(V); free(Wv); free(W); free(Fz);
}

/* rhsFreeSurface3D.f90:1-29 and rhsLinearFreeSurface3D.f90:1-13 (no applied pressure,
 * no breaking model, no pressure damping: P0 = Qr_x = Mr_t = NuD = Pd = 0). */
void owo_rhs(owo_tank *t, const double *E, const double *P, const double *W, double *rhsE, double *rhsP) {
    const grid_t *g = &t->g;
    double gr = t->p.g;
    (void)P;
    for (size_t q = 0; q < (size_t)g->nxy; q++) {
        if (!t->p.nonlinear) {
            rhsE[q] = W[q];
            rhsP[q] = -gr * E[q];
        } else if (g->ny == 1) {
            double ex = t->Ex[q], px = t->Px[q], w = W[q];
            rhsE[q] = -ex * px + w * (1.0 + ex * ex);
            rhsP[q] = -gr * E[q] - 0.5 * (px * px - w * w * (1.0 + ex * ex));
        } else {
            double ex = t->Ex[q], ey = t->Ey[q], px = t->Px[q], py = t->Py[q], w = W[q];
            rhsE[q] = -ex * px - ey * py + w * (1.0 + ex * ex + ey * ey);
            rhsP[q] = -gr * E[q] - 0.5 * (px * px + py * py - w * w * (1.0 + ex * ex + ey * ey));
        }
    }
}

/* Damping zones of RelaxationModule_new.f90:33-50 / 84-97 (WavegenONOFF = 0, PhiOnOff = 1,
 * no current): E <- gam E, P <- gam P over the zone, gam along the zone axis. */
void owo_add_relax_zone(owo_tank *t, int axis, int i0, int i1, int j0, int j1, const double *gam) {
    t->zones = realloc(t->zones, (size_t)(t->nzones + 1) * sizeof(zone_t));
    zone_t *z = &t->zones[t->nzones++];
    z->axis = axis; z->i0 = i0; z->i1 = i1; z->j0 = j0; z->j1 = j1;
    int len = axis == 0 ? i1 - i0 + 1 : j1 - j0 + 1;
    z->gam = xcalloc(len, sizeof(double));
    memcpy(z->gam, gam, len * sizeof(double));
}

static void relax(owo_tank *t, double *E, double *P) {
    const grid_t *g = &t->g;
    for (int zi = 0; zi < t->nzones; zi++) {
        zone_t *z = &t->zones[zi];
        for (int j = z->j0; j <= z->j1; j++)
            for (int i = z->i0; i <= z->i1; i++) {
                double gm = z->axis == 0 ? z->gam[i - z->i0] : z->gam[j - z->j0];
                E[I2(g, i, j)] = E[I2(g, i, j)] * gm;
                P[I2(g, i, j)] = P[I2(g, i, j)] * gm;
            }
    }
}

/* ------------------------------------------------------------------------- */
/* Multigrid                                                                  */
/* ------------------------------------------------------------------------- */
/* MGPreProcess.f90:449-458 (computed in single precision as there). */
static int no_grids(int nx, int ny, int nz) {
    int m = nx > ny ? nx : ny; m = m > nz ? m : nz;
    float tmp = (float)(log10f((float)(m - 1)) / log10(2.0));
    return (int)lroundf(tmp);
}

/* Low-order (2nd order, linearised) operator of a level, matrix-free.  Row semantics follow
 * BuildLinearSystemMatrix.f90 for the same grid, with the cross-derivative stencil applied on
 * the node's own transverse line (no quirk replication). */
static void low_apply(const level_t *L, const double *x, double *y) {
    const grid_t *g = &L->g;
    const double *ds = L->ds;
    size_t n = g->n;
    const int nxg = g->nxg, nyg = g->nyg, nzg = g->nzg;
    const int ra = g->ra, rb = g->rb, rc = g->rc;
    for (int j = 0; j < nyg; j++)
        for (int i = 0; i < nxg; i++)
            for (int k = 0; k < nzg; k++) {
                size_t q = I3(g, k, i, j);
                int wx = (g->gx && (i == 0 || i == nxg - 1));
                int wy = (g->gy && (j == 0 || j == nyg - 1));
                if ((wx && wy) || ((wx || wy) && k == 0)) { y[q] = x[q]; continue; }
                if (wx) {
                    int ib = i == 0 ? 1 : nxg - 2;
                    int ws = clampi(ib - g->a, 0, nxg - ra);
                    const double *S = g->sx + (ib * 2) * ra;
                    double s = 0.0;
                    for (int m = 0; m < ra; m++) s += S[m] * x[I3(g, k, ws + m, j)];
                    y[q] = s; continue;
                }
                if (wy) {
                    int jb = j == 0 ? 1 : nyg - 2;
                    int ws = clampi(jb - g->b, 0, nyg - rb);
                    const double *S = g->sy + (jb * 2) * rb;
                    double s = 0.0;
                    for (int m = 0; m < rb; m++) s += S[m] * x[I3(g, k, i, ws + m)];
                    y[q] = s; continue;
                }
                if (k == nzg - 1) { y[q] = x[q]; continue; }
                size_t p = I2(g, i, j);
                int wsx = clampi(i - g->a, 0, nxg - ra), wsy = clampi(j - g->b, 0, nyg - rb);
                if (k == 0) {
                    int kb = 1, wsz = clampi(kb - g->c, 0, nzg - rc);
                    const double *Sz = g->sz + (kb * 2) * rc;
                    double dz = 0.0, dx = 0.0, dy = 0.0;
                    for (int m = 0; m < rc; m++) dz += Sz[m] * x[I3(g, wsz + m, i, j)];
                    if (g->nx > 1) { const double *S = g->sx + (i * 2) * ra; for (int m = 0; m < ra; m++) dx += S[m] * x[I3(g, kb, wsx + m, j)]; }
                    if (g->ny > 1) { const double *S = g->sy + (j * 2) * rb; for (int m = 0; m < rb; m++) dy += S[m] * x[I3(g, kb, i, wsy + m)]; }
                    size_t qb = q + 1;
                    double hx = g->nx > 1 ? g->hx[p] : 0.0, hy = g->ny > 1 ? g->hy[p] : 0.0;
                    y[q] = (ds[4 * n + qb] + hx * ds[1 * n + qb] + hy * ds[3 * n + qb]) * dz + hx * dx + hy * dy;
                    continue;
                }
                int wsz = clampi(k - g->c, 0, nzg - rc);
                const double *Sz1 = g->sz + (k * 2) * rc, *Sz2 = Sz1 + rc;
                double dz = 0.0, dzz = 0.0, dxx = 0.0, dyy = 0.0, dxz = 0.0, dyz = 0.0;
                for (int m = 0; m < rc; m++) { dz += Sz1[m] * x[I3(g, wsz + m, i, j)]; dzz += Sz2[m] * x[I3(g, wsz + m, i, j)]; }
                if (g->nx > 1) {
                    const double *S2 = g->sx + (i * 2 + 1) * ra;
                    for (int m = 0; m < ra; m++) dxx += S2[m] * x[I3(g, k, wsx + m, j)];
                    for (int mz = 0; mz < rc; mz++) {
                        int kk = wsz + mz, sh = 0;
                        if (kk == 0) { if (wsx == 0) sh = 1; if (wsx + ra - 1 == nxg - 1) sh = -1; }
                        const double *S1 = g->sx + ((i - sh) * 2) * ra;
                        double dx = 0.0;
                        for (int m = 0; m < ra; m++) dx += S1[m] * x[I3(g, kk, wsx + sh + m, j)];
                        dxz += Sz1[mz] * dx;
                    }
                }
                if (g->ny > 1) {
                    const double *S2 = g->sy + (j * 2 + 1) * rb;
                    for (int m = 0; m < rb; m++) dyy += S2[m] * x[I3(g, k, i, wsy + m)];
                    for (int mz = 0; mz < rc; mz++) {
                        int kk = wsz + mz, sh = 0;
                        if (kk == 0) { if (wsy == 0) sh = 1; if (wsy + rb - 1 == nyg - 1) sh = -1; }
                        const double *S1 = g->sy + ((j - sh) * 2) * rb;
                        double dy = 0.0;
                        for (int m = 0; m < rb; m++) dy += S1[m] * x[I3(g, kk, i, wsy + sh + m)];
                        dyz += Sz1[mz] * dy;
                    }
                }
                double s2 = ds[1 * n + q], s3 = ds[2 * n + q], s4 = ds[3 * n + q], s5 = ds[4 * n + q];
                y[q] = dxx + dyy + s3 * dz + 2.0 * (s2 * dxz + s4 * dyz) + (s2 * s2 + s4 * s4 + s5 * s5) * dzz;
            }
}

/* COO row builder with SPARSKIT clncsr(job=1) semantics: duplicates summed into the first
 * occurrence, entries whose first occurrence is exactly zero are not inserted. */
typedef struct { int *col; double *val; int cnt, cap; } rowbuf_t;
static void rb_add(rowbuf_t *rb, int c, double v) {
    for (int q = 0; q < rb->cnt; q++) if (rb->col[q] == c) { rb->val[q] += v; return; }
    if (v == 0.0) return;
    if (rb->cnt == rb->cap) {
        rb->cap = rb->cap ? 2 * rb->cap : 64;
        rb->col = realloc(rb->col, rb->cap * sizeof(int));
        rb->val = realloc(rb->val, rb->cap * sizeof(double));
    }
    rb->col[rb->cnt] = c; rb->val[rb->cnt] = v; rb->cnt++;
}

/* Full-rank stencil box of node (k,i,j) (PrepareFullOperatorStencils.f90:179-211): clamped
 * windows in each direction; returns window starts. */
static void box(const grid_t *g, int k, int i, int j, int *wz, int *wx, int *wy) {
    *wz = clampi(k - g->c, 0, g->nzg - g->rc);
    *wx = clampi(i - g->a, 0, g->nxg - g->ra);
    *wy = clampi(j - g->b, 0, g->nyg - g->rb);
}

/* Assemble the level matrix in the reference's entry order (BuildLinearSystemMatrix.f90:37-324),
 * clean it (CleanSparseMatrixCOO.f90), convert to CSR and move the diagonal first
 * (CSRdiaREORDER.f90).  quirk != 0 reproduces the cross-derivative stencil replication of
 * ConstructTableCrossDerivatives.f90:312-347 (full-rank XZ stencil repeated over the y window,
 * YZ over the x window). */
static void assemble_csr(level_t *L, int quirk) {
    const grid_t *g = &L->g;
    size_t n = g->n;
    const int nxg = g->nxg, nyg = g->nyg, nzg = g->nzg, ra = g->ra, rb = g->rb, rc = g->rc;
    const double *ds = L->ds;
    rowbuf_t *rows = xcalloc(n, sizeof(rowbuf_t));
    /* Generic full-rank contributions: out row `row` gets coef * (stencil S applied along dir at node (k,i,j)). */
#define ADDX(row, coef, kk, ii, jj, S, ws) for (int m = 0; m < ra; m++) rb_add(&rows[row], (int)I3(g, kk, (ws) + m, jj), (coef) * (S)[m])
    /* 3D-or-2D interior and bottom rows, in the reference's loop order (j, i, k). */
    for (int j = g->gy; j < nyg - g->gy; j++)
        for (int i = g->gx; i < nxg - g->gx; i++)
            for (int k = 0; k < nzg; k++) {
                size_t q = I3(g, k, i, j), p = I2(g, i, j);
                if (k == nzg - 1) { rb_add(&rows[q], (int)q, 1.0); continue; }
                if (k == 0) {
                    /* bottom: full-rank box of node k=1; entries visited in box order (z fastest, then x, then y) */
                    int kb = 1, wz, wx, wy; box(g, kb, i, j, &wz, &wx, &wy);
                    size_t qb = q + 1;
                    double hx = g->nx > 1 ? g->hx[p] : 0.0, hy = g->ny > 1 ? g->hy[p] : 0.0;
                    double cz;
                    if (g->nx > 1 && g->ny > 1) cz = ds[4 * n + qb] + hx * ds[1 * n + qb] + hy * ds[3 * n + qb];
                    else cz = ds[4 * n + qb] + hx * ds[1 * n + qb];
                    const double *Sz = g->sz + (kb * 2) * rc, *Sx = g->sx + (i * 2) * ra, *Sy = g->sy + (j * 2) * rb;
                    for (int my = 0; my < rb; my++)
                        for (int mx = 0; mx < ra; mx++)
                            for (int mz = 0; mz < rc; mz++) {
                                int kk = wz + mz, ii = wx + mx, jj = wy + my;
                                /* the reference evaluates cz*Z + hx*X (+ hy*Y) per full-rank entry */
                                double zpart = (ii == i && jj == j) ? Sz[mz] : 0.0;
                                double xpart = (kk == kb && jj == j) ? Sx[mx] : 0.0;
                                double v;
                                if (g->ny > 1) {
                                    double ypart = (kk == kb && ii == i) ? Sy[my] : 0.0;
                                    v = cz * zpart + hx * xpart + hy * ypart;
                                } else {
                                    v = cz * zpart + hx * xpart;
                                }
                                rb_add(&rows[q], (int)I3(g, kk, ii, jj), v);
                            }
                    continue;
                }
                /* interior: XX + YY + s3 Z + (s2^2+s4^2+s5^2) ZZ over the box, then XZ, then YZ */
                int wz, wx, wy; box(g, k, i, j, &wz, &wx, &wy);
                double s2 = ds[1 * n + q], s3 = ds[2 * n + q], s4 = ds[3 * n + q], s5 = ds[4 * n + q];
                const double *Sz1 = g->sz + (k * 2) * rc, *Sz2 = Sz1 + rc;
                const double *Sx2 = g->sx + (i * 2 + 1) * ra, *Sy2 = g->sy + (j * 2 + 1) * rb;
                int d3 = (g->nx - 2 * 0 > 1 && g->ny > 1);
                for (int my = 0; my < rb; my++)
                    for (int mx = 0; mx < ra; mx++)
                        for (int mz = 0; mz < rc; mz++) {
                            int kk = wz + mz, ii = wx + mx, jj = wy + my;
                            double xx = (kk == k && jj == j && g->nx > 1) ? Sx2[mx] : 0.0;
                            double yy = (kk == k && ii == i && g->ny > 1) ? Sy2[my] : 0.0;
                            double z1 = (ii == i && jj == j) ? Sz1[mz] : 0.0;
                            double z2 = (ii == i && jj == j) ? Sz2[mz] : 0.0;
                            double v;
                            if (d3) v = xx + yy + s3 * z1 + (s2 * s2 + s4 * s4 + s5 * s5) * z2;
                            else v = xx + s3 * z1 + (s2 * s2 + s5 * s5) * z2;
                            rb_add(&rows[q], (int)I3(g, kk, ii, jj), v);
                        }
                /* XZ: sum over sigma window of Sz1 * (corner-corrected x stencil at that level) */
                if (g->nx > 1) {
                    int ny_rep = (quirk && g->ny > 1) ? rb : 1;
                    for (int my = 0; my < ny_rep; my++) {
                        int jj = (ny_rep == 1) ? j : wy + my;
                        for (int mx = 0; mx < ra; mx++)
                            for (int mz = 0; mz < rc; mz++) {
                                int kk = wz + mz, sh = 0;
                                if (kk == 0) { if (wx == 0) sh = 1; if (wx + ra - 1 == nxg - 1) sh = -1; }
                                const double *S1 = g->sx + ((i - sh) * 2) * ra;
                                rb_add(&rows[q], (int)I3(g, kk, wx + sh + mx, jj), 2.0 * (s2 * (S1[mx] * Sz1[mz])));
                            }
                    }
                }
                if (g->ny > 1) {
                    /* entry order of PACK(Index3D): sigma fastest, then x, then y */
                    int nx_rep = (quirk && g->nx > 1) ? ra : 1;
                    for (int my = 0; my < rb; my++)
                        for (int mx = 0; mx < nx_rep; mx++) {
                            int ii = (nx_rep == 1) ? i : wx + mx;
                            for (int mz = 0; mz < rc; mz++) {
                                int kk = wz + mz, sh = 0;
                                if (kk == 0) { if (wy == 0) sh = 1; if (wy + rb - 1 == nyg - 1) sh = -1; }
                                const double *S1 = g->sy + ((j - sh) * 2) * rb;
                                rb_add(&rows[q], (int)I3(g, kk, ii, wy + sh + my), 2.0 * (s4 * (S1[my] * Sz1[mz])));
                            }
                        }
                }
            }
    /* wall ghost rows k >= 1: x (or y) derivative at the adjacent boundary node */
    if (g->gx == 1)
        for (int side = 0; side < 2; side++) {
            int i = side == 0 ? 0 : nxg - 1, ib = side == 0 ? 1 : nxg - 2;
            for (int j = g->gy; j < nyg - g->gy; j++)
                for (int k = 1; k < nzg; k++) {
                    size_t q = I3(g, k, i, j);
                    int wz, wx, wy; box(g, k, ib, j, &wz, &wx, &wy);
                    const double *S = g->sx + (ib * 2) * ra;
                    for (int my = 0; my < rb; my++)
                        for (int mx = 0; mx < ra; mx++)
                            for (int mz = 0; mz < rc; mz++) {
                                int kk = wz + mz, jj = wy + my;
                                double v = (kk == k && jj == j) ? S[mx] : 0.0;
                                rb_add(&rows[q], (int)I3(g, kk, wx + mx, jj), v);
                            }
                }
        }
    if (g->gy == 1)
        for (int side = 0; side < 2; side++) {
            int j = side == 0 ? 0 : nyg - 1, jb = side == 0 ? 1 : nyg - 2;
            for (int i = g->gx; i < nxg - g->gx; i++)
                for (int k = 1; k < nzg; k++) {
                    size_t q = I3(g, k, i, j);
                    int wz, wx, wy; box(g, k, i, jb, &wz, &wx, &wy);
                    const double *S = g->sy + (jb * 2) * rb;
                    for (int my = 0; my < rb; my++)
                        for (int mx = 0; mx < ra; mx++)
                            for (int mz = 0; mz < rc; mz++) {
                                int kk = wz + mz, ii = wx + mx;
                                double v = (kk == k && ii == i) ? S[my] : 0.0;
                                rb_add(&rows[q], (int)I3(g, kk, ii, wy + my), v);
                            }
                }
        }
    /* identity rows: corner columns, bottom ghosts of wall columns */
    for (int j = 0; j < nyg; j++)
        for (int i = 0; i < nxg; i++) {
            int wx = (g->gx && (i == 0 || i == nxg - 1));
            int wy = (g->gy && (j == 0 || j == nyg - 1));
            if (wx && wy) { for (int k = 0; k < nzg; k++) rb_add(&rows[I3(g, k, i, j)], (int)I3(g, k, i, j), 1.0); }
            else if (wx || wy) rb_add(&rows[I3(g, 0, i, j)], (int)I3(g, 0, i, j), 1.0);
        }
#undef ADDX
    /* CSR with diagonal first */
    int nnz = 0;
    for (size_t r = 0; r < n; r++) nnz += rows[r].cnt;
    L->nnz = nnz;
    L->rp = xcalloc(n + 1, sizeof(int));
    L->ci = xcalloc(nnz, sizeof(int));
    L->va = xcalloc(nnz, sizeof(double));
    int pos = 0;
    for (size_t r = 0; r < n; r++) {
        L->rp[r] = pos;
        rowbuf_t *rbf = &rows[r];
        for (int q = 0; q < rbf->cnt; q++) { L->ci[pos + q] = rbf->col[q]; L->va[pos + q] = rbf->val[q]; }
        for (int q = 0; q < rbf->cnt; q++)
            if (rbf->col[q] == (int)r) {
                int tc = L->ci[pos]; double tv = L->va[pos];
                L->ci[pos] = (int)r; L->va[pos] = rbf->val[q];
                L->ci[pos + q] = tc; L->va[pos + q] = tv;
                break;
            }
        pos += rbf->cnt;
        free(rbf->col); free(rbf->val);
    }
    L->rp[n] = pos;
    free(rows);
}

/* GaussSeidel.f90: forward lexicographic sweep, diagonal stored first. */
static void gs_csr(const level_t *L, double *x, const double *b) {
    size_t n = L->g.n;
    for (size_t i = 0; i < n; i++) {
        double s = 0.0;
        for (int q = L->rp[i] + 1; q < L->rp[i + 1]; q++) s += x[L->ci[q]] * L->va[q];
        x[i] = (b[i] - s) / L->va[L->rp[i]];
    }
}

/* GaussSeidelSpecialPoints.f90 */
static void gs_special(const level_t *L, double *x, const double *b) {
    for (int t = 0; t < L->nghost; t++) {
        int i = L->ighost[t];
        double s = 0.0;
        for (int q = L->rp[i] + 1; q < L->rp[i + 1]; q++) s += x[L->ci[q]] * L->va[q];
        x[i] = (b[i] - s) / L->va[L->rp[i]];
    }
}

/* amux (SPARSKIT2 CSR product) */
static void amux_csr(const level_t *L, const double *x, double *y) {
    size_t n = L->g.n;
    for (size_t i = 0; i < n; i++) {
        double s = 0.0;
        for (int q = L->rp[i]; q < L->rp[i + 1]; q++) s += L->va[q] * x[L->ci[q]];
        y[i] = s;
    }
}

/* Zebra z-line smoother (the GPU preconditioner's smoother): columns of one colour
 * ((i+j) mod 2) are relaxed by an exact tridiagonal solve in sigma with the neighbouring
 * columns frozen; afterwards the wall ghost points are relaxed pointwise. */
static void zebra_colour(const level_t *L, double *x, const double *b, int colour);
static void ghost_relax(const level_t *L, double *x, const double *b);

/* MGRestriction3DExplicit.f90:1-127 */
static void restrict_ref(const level_t *F, const level_t *Cl, const double *rf, double *rc, double *tmp) {
    const grid_t *f = &F->g, *c = &Cl->g;
    int xc = Cl->xc, yc = Cl->yc, zc = Cl->zc;
    memcpy(tmp, rf, f->n * sizeof(double));
    memset(rc, 0, c->n * sizeof(double));
    int nxf = f->nx, nyf = f->ny, nzf = f->nz, gx = f->gx, gy = f->gy;
    if (xc) {
        for (int i = gx; i <= nxf + gx - 1; i += 2) {
            double dW = f->x[i] - f->x[i - 1], dE = f->x[i + 1] - f->x[i];
            double dxc = 1.0 / (dW + dE);
            for (int k = 1; k <= nzf - 1; k++)
                for (int j = gy; j <= nyf + gy - 1; j++)
                    tmp[I3(f, k, i, j)] = (0.5 * dxc) * (dW * tmp[I3(f, k, i - 1, j)] + (dW + dE) * tmp[I3(f, k, i, j)] + dE * tmp[I3(f, k, i + 1, j)]);
        }
    }
    if (yc) {
        for (int j = gy; j <= nyf + gy - 1; j += 2) {
            double dS = f->y[j] - f->y[j - 1], dN = f->y[j + 1] - f->y[j];
            double dyc = 1.0 / (dS + dN);
            for (int k = 1; k <= nzf - 1; k++)
                for (int i = gx; i <= nxf + gx - 1; i += 1 + xc)
                    tmp[I3(f, k, i, j)] = (0.5 * dyc) * (dS * tmp[I3(f, k, i, j - 1)] + (dN + dS) * tmp[I3(f, k, i, j)] + dN * tmp[I3(f, k, i, j + 1)]);
        }
    }
    if (zc) {
        for (int k = 1; k <= nzf - 1; k += 2) {
            double dU = f->z[k + 1] - f->z[k], dD = f->z[k] - f->z[k - 1];
            double dzc = 1.0 / (dD + dU);
            for (int i = gx; i <= nxf + gx - 1; i += 1 + xc)
                for (int j = gy; j <= nyf + gy - 1; j += 1 + yc)
                    tmp[I3(f, k, i, j)] = (0.5 * dzc) * (dD * tmp[I3(f, k - 1, i, j)] + dU * tmp[I3(f, k + 1, i, j)]) + 0.5 * tmp[I3(f, k, i, j)];
        }
    }
    for (int jc = 0; jc < c->ny; jc++)
        for (int ic = 0; ic < c->nx; ic++)
            for (int kc = 0; kc < c->nz; kc++)
                rc[I3(c, 1 + kc, c->gx + ic, c->gy + jc)] = tmp[I3(f, 1 + kc * (1 + zc), gx + ic * (1 + xc), gy + jc * (1 + yc))];
}

/* MGProlongation3DExplicit.f90:1-108 */
static void prolong_ref(const level_t *F, const level_t *Cl, const double *ec, double *ef) {
    const grid_t *f = &F->g, *c = &Cl->g;
    int XF = Cl->xc ? 2 : 1, YF = Cl->yc ? 2 : 1, ZF = Cl->zc ? 2 : 1;
    int nxf = f->nx, nyf = f->ny, nzf = f->nz, gx = f->gx, gy = f->gy;
    memset(ef, 0, f->n * sizeof(double));
    /* injection at coincident nodes, free surface excluded */
    for (int jc = 0; jc < c->ny; jc++)
        for (int ic = 0; ic < c->nx; ic++)
            for (int kc = 0; kc < c->nz - 1; kc++)
                ef[I3(f, 1 + kc * ZF, gx + ic * XF, gy + jc * YF)] = ec[I3(c, 1 + kc, c->gx + ic, c->gy + jc)];
    if (XF == 2)
        for (int j = gy; j <= nyf + gy - 1; j += YF)
            for (int i = 1 + gx; i <= nxf + gx - 1; i += XF) {
                double dW = f->x[i] - f->x[i - 1], dE = f->x[i + 1] - f->x[i];
                double dxf = 1.0 / (dW + dE);
                for (int k = 1; k <= nzf - 1; k += ZF)
                    ef[I3(f, k, i, j)] = (dE * ef[I3(f, k, i - 1, j)] + dW * ef[I3(f, k, i + 1, j)]) * dxf;
            }
    if (YF == 2)
        for (int i = gx; i <= nxf + gx - 1; i += 1)
            for (int j = 1 + gy; j <= nyf + gy - 1; j += YF) {
                double dS = f->y[j] - f->y[j - 1], dN = f->y[j + 1] - f->y[j];
                double dyf = 1.0 / (dS + dN);
                for (int k = 1; k <= nzf - 1; k += ZF)
                    ef[I3(f, k, i, j)] = (dN * ef[I3(f, k, i, j - 1)] + dS * ef[I3(f, k, i, j + 1)]) * dyf;
            }
    if (ZF == 2)
        for (int k = 2; k <= nzf - 1; k += ZF) {
            double dD = f->z[k] - f->z[k - 1], dU = f->z[k + 1] - f->z[k];
            double dzf = 1.0 / (dD + dU);
            for (int j = gy; j <= nyf + gy - 1; j += 1)
                for (int i = gx; i <= nxf + gx - 1; i += 1)
                    ef[I3(f, k, i, j)] = (dU * ef[I3(f, k - 1, i, j)] + dD * ef[I3(f, k + 1, i, j)]) * dzf;
        }
}

/* Dense LU with partial pivoting (exact coarse-level solve standing in for MA41). */
static void dense_lu(double *A, int *piv, int n) {
    for (int c = 0; c < n; c++) {
        int pr = c; double best = fabs(A[(size_t)c * n + c]);
        for (int r = c + 1; r < n; r++) if (fabs(A[(size_t)r * n + c]) > best) { best = fabs(A[(size_t)r * n + c]); pr = r; }
        piv[c] = pr;
        if (pr != c) for (int q = 0; q < n; q++) { double t = A[(size_t)c * n + q]; A[(size_t)c * n + q] = A[(size_t)pr * n + q]; A[(size_t)pr * n + q] = t; }
        double d = A[(size_t)c * n + c];
        for (int r = c + 1; r < n; r++) {
            double f = A[(size_t)r * n + c] / d;
            A[(size_t)r * n + c] = f;
            for (int q = c + 1; q < n; q++) A[(size_t)r * n + q] -= f * A[(size_t)c * n + q];
        }
    }
}
static void dense_solve(const double *A, const int *piv, int n, double *b) {
    for (int c = 0; c < n; c++) { if (piv[c] != c) { double t = b[c]; b[c] = b[piv[c]]; b[piv[c]] = t; } }
    for (int r = 0; r < n; r++) { double s = b[r]; for (int q = 0; q < r; q++) s -= A[(size_t)r * n + q] * b[q]; b[r] = s; }
    for (int r = n - 1; r >= 0; r--) { double s = b[r]; for (int q = r + 1; q < n; q++) s -= A[(size_t)r * n + q] * b[q]; b[r] = s / A[(size_t)r * n + r]; }
}

static void level_residual(const owo_tank *t, const level_t *L, const double *x, const double *b, double *r) {
    if (t->p.mg_mode == 0) amux_csr(L, x, r); else low_apply(L, x, r);
    for (size_t i = 0; i < L->g.n; i++) r[i] = b[i] - r[i];
}

static void smooth_sweeps(const owo_tank *t, const level_t *L, double *x, const double *b, int nu, int initial_ghost) {
    if (t->p.mg_mode == 0) {
        if (initial_ghost) gs_special(L, x, b);
        for (int s = 0; s < nu; s++) { gs_csr(L, x, b); gs_special(L, x, b); }
    } else {
        if (initial_ghost) ghost_relax(L, x, b);
        for (int s = 0; s < nu; s++) { zebra_colour(L, x, b, 0); zebra_colour(L, x, b, 1); ghost_relax(L, x, b); }
    }
}

/* MultiGridSolver.f90:5-192 (V cycle, gam = 1) */
static void vcycle(owo_tank *t, int k, double *x, const double *b) {
    level_t *L = &t->lev[k];
    if (k == 0) {
        memcpy(x, b, L->g.n * sizeof(double));
        dense_solve(L->lu, L->piv, (int)L->g.n, x);
        return;
    }
    level_t *C = &t->lev[k - 1];
    smooth_sweeps(t, L, x, b, t->p.nu1, 1);
    level_residual(t, L, x, b, L->r);
    restrict_ref(L, C, L->r, C->b, L->t1);
    memset(C->x, 0, C->g.n * sizeof(double));
    vcycle(t, k - 1, C->x, C->b);
    prolong_ref(L, C, C->x, L->t1);
    for (size_t i = 0; i < L->g.n; i++) x[i] = x[i] + L->t1[i];
    smooth_sweeps(t, L, x, b, t->p.nu2, 1);
}

/* MultiGridSolveLinearSystem.f90:26-91 */
void owo_precond(owo_tank *t, const double *r, double *u) {
    size_t n = t->g.n;
    if (t->p.precond != 3) { memcpy(u, r, n * sizeof(double)); return; }
    memset(u, 0, n * sizeof(double));
    for (int it = 0; it < t->p.maxit; it++) {
        if (t->nlev == 1 && it > 0) break;
        vcycle(t, t->nlev - 1, u, r);
    }
}

static void zebra_colour(const level_t *L, double *x, const double *b, int colour) {
    const grid_t *g = &L->g;
    size_t n = g->n;
    const int nxg = g->nxg, nyg = g->nyg, nzg = g->nzg, ra = g->ra, rb = g->rb, rc = g->rc;
    const double *ds = L->ds;
    double lo[64], di[64], up[64], rhs[64], e2 = 0.0;
    for (int j = g->gy; j < nyg - g->gy; j++)
        for (int i = g->gx; i < nxg - g->gx; i++) {
            if (((i + j) & 1) != colour) continue;
            size_t p = I2(g, i, j);
            int wsx = clampi(i - g->a, 0, nxg - ra), wsy = clampi(j - g->b, 0, nyg - rb);
            /* own-column weights of x / y first-derivative stencils per level (CD shift at k=0) */
            for (int k = 0; k < nzg; k++) { lo[k] = di[k] = up[k] = 0.0; }
            e2 = 0.0;
            for (int k = 0; k < nzg; k++) {
                size_t q = I3(g, k, i, j);
                double off = 0.0;
                if (k == nzg - 1) { di[k] = 1.0; rhs[k] = b[q]; continue; }
                if (k == 0) {
                    int kb = 1, wsz = clampi(kb - g->c, 0, nzg - rc);
                    const double *Sz = g->sz + (kb * 2) * rc;
                    size_t qb = q + 1;
                    double hx = g->nx > 1 ? g->hx[p] : 0.0, hy = g->ny > 1 ? g->hy[p] : 0.0;
                    double cz = ds[4 * n + qb] + hx * ds[1 * n + qb] + hy * ds[3 * n + qb];
                    double a[3] = {0.0, 0.0, 0.0};
                    for (int m = 0; m < rc; m++) a[wsz + m] += cz * Sz[m];
                    if (g->nx > 1) { const double *S = g->sx + (i * 2) * ra; for (int m = 0; m < ra; m++) { if (wsx + m == i) a[kb] += hx * S[m]; else off += hx * S[m] * x[I3(g, kb, wsx + m, j)]; } }
                    if (g->ny > 1) { const double *S = g->sy + (j * 2) * rb; for (int m = 0; m < rb; m++) { if (wsy + m == j) a[kb] += hy * S[m]; else off += hy * S[m] * x[I3(g, kb, i, wsy + m)]; } }
                    di[0] = a[0]; up[0] = a[1]; e2 = a[2];
                    rhs[0] = b[q] - off;
                    continue;
                }
                int wsz = clampi(k - g->c, 0, nzg - rc);
                const double *Sz1 = g->sz + (k * 2) * rc, *Sz2 = Sz1 + rc;
                double s2 = ds[1 * n + q], s3 = ds[2 * n + q], s4 = ds[3 * n + q], s5 = ds[4 * n + q];
                double czz = s2 * s2 + s4 * s4 + s5 * s5;
                double a[3] = {0.0, 0.0, 0.0}; /* coefficients of x[k-1], x[k], x[k+1] */
                for (int m = 0; m < rc; m++) a[wsz + m - (k - 1)] += s3 * Sz1[m] + czz * Sz2[m];
                if (g->nx > 1) {
                    const double *S2 = g->sx + (i * 2 + 1) * ra;
                    for (int m = 0; m < ra; m++) { if (wsx + m == i) a[1] += S2[m]; else off += S2[m] * x[I3(g, k, wsx + m, j)]; }
                    for (int mz = 0; mz < rc; mz++) {
                        int kk = wsz + mz, sh = 0;
                        if (kk == 0) { if (wsx == 0) sh = 1; if (wsx + ra - 1 == nxg - 1) sh = -1; }
                        const double *S1 = g->sx + ((i - sh) * 2) * ra;
                        for (int m = 0; m < ra; m++) {
                            double w = 2.0 * s2 * Sz1[mz] * S1[m];
                            int ii = wsx + sh + m;
                            if (ii == i) a[kk - (k - 1)] += w;
                            else if (ii - i != 2 && i - ii != 2) off += w * x[I3(g, kk, ii, j)]; /* same colour: uncoupled */
                        }
                    }
                }
                if (g->ny > 1) {
                    const double *S2 = g->sy + (j * 2 + 1) * rb;
                    for (int m = 0; m < rb; m++) { if (wsy + m == j) a[1] += S2[m]; else off += S2[m] * x[I3(g, k, i, wsy + m)]; }
                    for (int mz = 0; mz < rc; mz++) {
                        int kk = wsz + mz, sh = 0;
                        if (kk == 0) { if (wsy == 0) sh = 1; if (wsy + rb - 1 == nyg - 1) sh = -1; }
                        const double *S1 = g->sy + ((j - sh) * 2) * rb;
                        for (int m = 0; m < rb; m++) {
                            double w = 2.0 * s4 * Sz1[mz] * S1[m];
                            int jj = wsy + sh + m;
                            if (jj == j) a[kk - (k - 1)] += w;
                            else if (jj - j != 2 && j - jj != 2) off += w * x[I3(g, kk, i, jj)]; /* same colour: uncoupled */
                        }
                    }
                }
                lo[k] = a[0]; di[k] = a[1]; up[k] = a[2];
                rhs[k] = b[q] - off;
            }
            /* eliminate the bottom row's x[2] entry with row 1, then Thomas */
            if (e2 != 0.0) {
                double f = e2 / up[1];
                di[0] -= f * lo[1];
                up[0] -= f * di[1];
                rhs[0] -= f * rhs[1];
            }
            for (int k = 1; k < nzg; k++) {
                double f = lo[k] / di[k - 1];
                di[k] -= f * up[k - 1];
                rhs[k] -= f * rhs[k - 1];
            }
            double *xc = x + I3(g, 0, i, j);
            xc[nzg - 1] = rhs[nzg - 1] / di[nzg - 1];
            for (int k = nzg - 2; k >= 0; k--) xc[k] = (rhs[k] - up[k] * xc[k + 1]) / di[k];
        }
}

static void ghost_relax(const level_t *L, double *x, const double *b) {
    const grid_t *g = &L->g;
    const int nxg = g->nxg, nyg = g->nyg, nzg = g->nzg, ra = g->ra, rb = g->rb;
    for (int j = 0; j < nyg; j++)
        for (int i = 0; i < nxg; i++) {
            int wx = (g->gx && (i == 0 || i == nxg - 1));
            int wy = (g->gy && (j == 0 || j == nyg - 1));
            if (!wx && !wy) continue;
            for (int k = 0; k < nzg; k++) {
                size_t q = I3(g, k, i, j);
                if ((wx && wy) || k == 0) { x[q] = b[q]; continue; }
                if (wx) {
                    int ib = i == 0 ? 1 : nxg - 2, ws = clampi(ib - g->a, 0, nxg - ra);
                    const double *S = g->sx + (ib * 2) * ra;
                    double s = 0.0, dg = 0.0;
                    for (int m = 0; m < ra; m++) { if (ws + m == i) dg = S[m]; else s += S[m] * x[I3(g, k, ws + m, j)]; }
                    x[q] = (b[q] - s) / dg;
                } else {
                    int jb = j == 0 ? 1 : nyg - 2, ws = clampi(jb - g->b, 0, nyg - rb);
                    const double *S = g->sy + (jb * 2) * rb;
                    double s = 0.0, dg = 0.0;
                    for (int m = 0; m < rb; m++) { if (ws + m == j) dg = S[m]; else s += S[m] * x[I3(g, k, i, ws + m)]; }
                    x[q] = (b[q] - s) / dg;
                }
            }
        }
}

/* Directions halved from an (nx, ny, nz) level: coarsening strategy 0 (MGPreProcess.f90:140-178).  Where
 * it would halve an even column count (the reference stops with an error, MGPreProcess.f90:183-211) the
 * column is kept and x / y are halved while they have an odd count above 3 (the GPU build's rule for
 * even Nz).  Returns 0 when nothing can be halved. */
static int coarsen_flags(int nx, int ny, int nz, int *xc, int *yc, int *zc) {
    *xc = *yc = *zc = 0;
    if (nx == 1) { if (ny == nz) { *yc = 1; *zc = 1; } else *yc = 1; }
    else if (ny == 1) { if (nx == nz) { *xc = 1; *zc = 1; } else *xc = 1; }
    else if (nx == ny && nx == nz) { *xc = *yc = *zc = 1; }
    else { if (nx > nz) *xc = 1; if (ny > nz) *yc = 1; if (nx < nz && ny < nz) *zc = 1; }
    if (*zc && nz % 2 == 0) {
        *zc = 0;
        *xc = (nx > 3 && nx % 2 == 1);
        *yc = (ny > 3 && ny % 2 == 1);
    }
    return *xc || *yc || *zc;
}

/* MGPreProcess.f90:30-318 (coarsening strategy 0) and 321-439 (operators, ghost maps). */
static int mg_setup(owo_tank *t) {
    const grid_t *fg = &t->g;
    int L = no_grids(fg->nx, fg->ny, fg->nz);
    if (t->p.mg_maxgrids > 0 && L > t->p.mg_maxgrids) L = t->p.mg_maxgrids;
    if (L < 1) L = 1;
    {   /* stop where nothing can be halved */
        int nx = fg->nx, ny = fg->ny, nz = fg->nz, xc, yc, zc;
        for (int q = 1; q < L; q++) {
            if (!coarsen_flags(nx, ny, nz, &xc, &yc, &zc)) { L = q; break; }
            if (xc) nx = (nx - 1) / 2 + 1;
            if (yc) ny = (ny - 1) / 2 + 1;
            if (zc) nz = (nz - 1) / 2 + 1;
        }
    }
    t->nlev = L;
    t->lev = xcalloc(L, sizeof(level_t));
    /* finest level: fine grid with 2nd-order stencils */
    for (int k = L - 1; k >= 0; k--) {
        level_t *lv = &t->lev[k];
        grid_t *g = &lv->g;
        if (k == L - 1) {
            g->nx = fg->nx; g->ny = fg->ny; g->nz = fg->nz;
        } else {
            const grid_t *f = &t->lev[k + 1].g;
            int NFx = f->nx, NFy = f->ny, NFz = f->nz, xc = 0, yc = 0, zc = 0;
            coarsen_flags(NFx, NFy, NFz, &xc, &yc, &zc);
            if (xc && NFx % 2 == 0) { fprintf(stderr, "ow3d_oracle: multigrid needs odd Nx (got %d)\n", NFx); return -1; }
            if (yc && NFy % 2 == 0) { fprintf(stderr, "ow3d_oracle: multigrid needs odd Ny (got %d)\n", NFy); return -1; }
            g->nx = xc ? (NFx - 1) / 2 + 1 : NFx;
            g->ny = yc ? (NFy - 1) / 2 + 1 : NFy;
            g->nz = zc ? (NFz - 1) / 2 + 1 : NFz;
            if (zc && g->nz % 2 == 0) { fprintf(stderr, "ow3d_oracle: multigrid needs odd coarse Nz\n"); return -1; }
            lv->xc = xc; lv->yc = yc; lv->zc = zc;
        }
        g->gx = fg->gx; g->gy = fg->gy;
        g->a = g->nx > 1 ? 1 : 0; g->b = g->ny > 1 ? 1 : 0; g->c = 1;
        grid_alloc(g);
        if (k == L - 1) {
            memcpy(g->x, fg->x, g->nxg * sizeof(double));
            memcpy(g->y, fg->y, g->nyg * sizeof(double));
            memcpy(g->z, fg->z, g->nzg * sizeof(double));
            memcpy(g->h, fg->h, g->nxy * sizeof(double));
            memcpy(g->hx, fg->hx, g->nxy * sizeof(double));
            memcpy(g->hy, fg->hy, g->nxy * sizeof(double));
            memcpy(g->hxx, fg->hxx, g->nxy * sizeof(double));
            memcpy(g->hyy, fg->hyy, g->nxy * sizeof(double));
        } else {
            const grid_t *f = &t->lev[k + 1].g;
            int XF = lv->xc ? 2 : 1, YF = lv->yc ? 2 : 1, ZF = lv->zc ? 2 : 1;
            if (g->nx > 1) {
                for (int ic = 0; ic < g->nx; ic++) g->x[g->gx + ic] = f->x[f->gx + ic * XF];
                if (g->gx) {
                    g->x[0] = g->x[1] - (g->x[2] - g->x[1]);
                    g->x[g->nxg - 1] = g->x[g->nxg - 2] + (g->x[g->nxg - 2] - g->x[g->nxg - 3]);
                }
            } else g->x[0] = f->x[0];
            if (g->ny > 1) {
                for (int jc = 0; jc < g->ny; jc++) g->y[g->gy + jc] = f->y[f->gy + jc * YF];
                if (g->gy) {
                    g->y[0] = g->y[1] - (g->y[2] - g->y[1]);
                    g->y[g->nyg - 1] = g->y[g->nyg - 2] + (g->y[g->nyg - 2] - g->y[g->nyg - 3]);
                }
            } else g->y[0] = f->y[0];
            g->z[0] = -f->z[1 + ZF];
            for (int kc = 0; kc < g->nz; kc++) g->z[1 + kc] = f->z[1 + kc * ZF];
            for (int jc = 0; jc < g->ny; jc++)
                for (int ic = 0; ic < g->nx; ic++) {
                    size_t pc = I2(g, g->gx + ic, g->gy + jc), pf = I2(f, f->gx + ic * XF, f->gy + jc * YF);
                    g->h[pc] = f->h[pf]; g->hx[pc] = f->hx[pf]; g->hxx[pc] = f->hxx[pf];
                    g->hy[pc] = f->hy[pf]; g->hyy[pc] = f->hyy[pf];
                }
            for (int j = 0; j < g->nyg; j++) { g->h[I2(g, 0, j)] = g->h[I2(g, g->gx, j)]; g->h[I2(g, g->nxg - 1, j)] = g->h[I2(g, g->nxg - 1 - g->gx, j)]; }
            for (int i = 0; i < g->nxg; i++) { g->h[I2(g, i, 0)] = g->h[I2(g, i, g->gy)]; g->h[I2(g, i, g->nyg - 1)] = g->h[I2(g, i, g->nyg - 1 - g->gy)]; }
        }
    }
    for (int k = 0; k < L; k++) {
        level_t *lv = &t->lev[k];
        grid_t *g = &lv->g;
        grid_stencils(g);
        lv->ds = xcalloc(5 * (size_t)g->n, sizeof(double));
        compute_dsigma(g, NULL, NULL, NULL, NULL, NULL, lv->ds);
        lv->x = xcalloc(g->n, sizeof(double)); lv->b = xcalloc(g->n, sizeof(double));
        lv->r = xcalloc(g->n, sizeof(double)); lv->t1 = xcalloc(g->n, sizeof(double));
        if (t->p.mg_mode == 0 || k == 0) assemble_csr(lv, t->p.mg_mode == 0 ? t->p.cd_quirk : 0);
        if (k == 0) {
            int n = (int)g->n;
            lv->lu = xcalloc((size_t)n * n, sizeof(double));
            lv->piv = xcalloc(n, sizeof(int));
            if (t->p.mg_mode == 0) {
                for (int r = 0; r < n; r++) for (int q = lv->rp[r]; q < lv->rp[r + 1]; q++) lv->lu[(size_t)r * n + lv->ci[q]] += lv->va[q];
            } else {
                double *e = xcalloc(n, sizeof(double)), *col = xcalloc(n, sizeof(double));
                for (int c = 0; c < n; c++) { e[c] = 1.0; low_apply(lv, e, col); e[c] = 0.0; for (int r = 0; r < n; r++) lv->lu[(size_t)r * n + c] = col[r]; }
                free(e); free(col);
            }
            dense_lu(lv->lu, lv->piv, n);
        }
        /* special points: bottom ghost layer and wall ghost columns, in storage order */
        int cnt = 0;
        lv->ighost = xcalloc(g->n, sizeof(int));
        for (int j = 0; j < g->nyg; j++)
            for (int i = 0; i < g->nxg; i++)
                for (int kk = 0; kk < g->nzg; kk++) {
                    int wx = g->gx && (i == 0 || i == g->nxg - 1), wy = g->gy && (j == 0 || j == g->nyg - 1);
                    if (kk == 0 || wx || wy) lv->ighost[cnt++] = (int)I3(g, kk, i, j);
                }
        lv->nghost = cnt;
    }
    return 0;
}

int owo_mg_level_dims(const owo_tank *t, int level, int *out3) {
    if (level < 0 || level >= t->nlev) return -1;
    const grid_t *g = &t->lev[level].g;
    out3[0] = g->nxg; out3[1] = g->nyg; out3[2] = g->nzg;
    return 0;
}

void owo_lowmatvec(owo_tank *t, int level, const double *phi, double *out) {
    level_t *L = &t->lev[level];
    if (L->rp && t->p.mg_mode == 0) amux_csr(L, phi, out); else low_apply(L, phi, out);
}

void owo_restrict(owo_tank *t, int fine_level, const double *rf, double *rc) {
    restrict_ref(&t->lev[fine_level], &t->lev[fine_level - 1], rf, rc, t->lev[fine_level].t1);
}
void owo_prolong(owo_tank *t, int fine_level, const double *ec, double *ef) {
    prolong_ref(&t->lev[fine_level], &t->lev[fine_level - 1], ec, ef);
}
void owo_smooth(owo_tank *t, int level, const double *b, double *x, int sweeps) {
    smooth_sweeps(t, &t->lev[level], x, b, sweeps, 1);
}
/* Level grid as built by mg_setup: dims6 = nx, ny, nz, nxg, nyg, nzg; coordinate and depth arrays. */
int owo_level_grid(const owo_tank *t, int level, int *dims6, double *x, double *y, double *z, double *h,
                   double *hx, double *hy, double *hxx, double *hyy) {
    if (level < 0 || level >= t->nlev) return -1;
    const grid_t *g = &t->lev[level].g;
    dims6[0] = g->nx; dims6[1] = g->ny; dims6[2] = g->nz; dims6[3] = g->nxg; dims6[4] = g->nyg; dims6[5] = g->nzg;
    if (x) memcpy(x, g->x, g->nxg * 8);
    if (y) memcpy(y, g->y, g->nyg * 8);
    if (z) memcpy(z, g->z, g->nzg * 8);
    if (h) memcpy(h, g->h, g->nxy * 8);
    if (hx) memcpy(hx, g->hx, g->nxy * 8);
    if (hy) memcpy(hy, g->hy, g->nxy * 8);
    if (hxx) memcpy(hxx, g->hxx, g->nxy * 8);
    if (hyy) memcpy(hyy, g->hyy, g->nxy * 8);
    return 0;
}
int owo_level_ghosts(const owo_tank *t, int level, int *out) {
    const level_t *L = &t->lev[level];
    if (out) memcpy(out, L->ighost, L->nghost * sizeof(int));
    return L->nghost;
}
/* One GaussSeidel sweep followed by one GaussSeidelSpecialPoints pass (reference mode). */
void owo_gs_once(owo_tank *t, int level, const double *b, double *x) {
    gs_csr(&t->lev[level], x, b);
    gs_special(&t->lev[level], x, b);
}
/* One colour of the zebra-line sweep, or (colour 2) the wall-ghost pass, on its own. */
void owo_zebra(owo_tank *t, int level, const double *b, double *x, int colour) {
    if (colour == 2) ghost_relax(&t->lev[level], x, b);
    else zebra_colour(&t->lev[level], x, b, colour);
}
int owo_csr(owo_tank *t, int level, int *nrow, int *nnz, const int **rowptr, const int **col, const double **val) {
    level_t *L = &t->lev[level];
    if (!L->rp) return -1;
    *nrow = (int)L->g.n; *nnz = L->nnz; *rowptr = L->rp; *col = L->ci; *val = L->va;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Krylov drivers                                                            */
/* ------------------------------------------------------------------------- */
static double dot(const double *a, const double *b, size_t n) {
    double s = 0.0;
    for (size_t i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* Left-preconditioned GMRES(m), one cycle, as driven by iterative_solution.f90:19-56 through
 * SPARSKIT2's `gmres` (not part of the reference tree; restated from Saad & Schultz 1986 and
 * SPARSKIT2's documented protocol): Arnoldi on M^-1 A with modified Gram-Schmidt plus
 * selective re-orthogonalisation, Givens rotations, stop when the estimated preconditioned
 * residual <= reltol*||M^-1 r0|| + abstol, at most m matrix-vector products in total (the
 * initial residual counts as one).  Returns 0 when converged, -1 when the product budget ran out.
 * gmres_side 1: flexible right preconditioning (FGMRES, Saad, SIAM J. Sci. Comput. 14, 1993): Arnoldi on
 * A z_j with z_j = M^-1 v_j kept, x = x0 + Z y, stop on the estimated true residual <= reltol*||r0|| +
 * abstol (no counterpart in the reference; the product uses it with an FP32 multigrid, which is linear
 * only to FP32 rounding). */
static int gmres_solve(owo_tank *t, const double *rhs, double *sol, int *iters, double *res) {
    size_t n = t->g.n;
    int m = t->p.gmres_m;
    if (m <= 1) m = 15;
    const int flex = t->p.gmres_side == 1;
    double *V = t->kw, *w = t->v1, *z = t->v2, *Z = t->kz;
    double *H = xcalloc((size_t)(m + 1) * (m + 1), sizeof(double));
    double *cs = xcalloc(m + 1, sizeof(double)), *sn = xcalloc(m + 1, sizeof(double)), *gv = xcalloc(m + 2, sizeof(double));
    int matvecs = 0, k = 0, status = -1;
    owo_matvec(t, sol, w); matvecs++;
    for (size_t i = 0; i < n; i++) w[i] = rhs[i] - w[i];
    if (flex) memcpy(V, w, n * sizeof(double));
    else owo_precond(t, w, V);
    double beta = sqrt(dot(V, V, n));
    double tol = t->p.reltol * beta + t->p.abstol;
    *res = beta;
    if (beta <= tol) { status = 0; goto done; }
    for (size_t i = 0; i < n; i++) V[i] *= 1.0 / beta;
    gv[0] = beta;
    double est = beta;
    while (1) {
        k++;
        double *vk = V + (size_t)(k - 1) * n, *vn = V + (size_t)k * n;
        if (flex) {
            double *zk = Z + (size_t)(k - 1) * n;
            owo_precond(t, vk, zk);
            owo_matvec(t, zk, vn); matvecs++;
        } else {
            owo_matvec(t, vk, w); matvecs++;
            owo_precond(t, w, vn);
        }
        /* MGS with selective re-orthogonalisation (cosine > 0.99) */
        double nrm0 = dot(vn, vn, n), thr = nrm0 * 0.99 * 0.99;
        for (int i = 0; i < k; i++) {
            double *vi = V + (size_t)i * n;
            double hij = dot(vi, vn, n);
            H[(size_t)i * m + (k - 1)] = hij;
            for (size_t q = 0; q < n; q++) vn[q] -= hij * vi[q];
            if (hij * hij > thr) {
                double h2 = dot(vi, vn, n);
                H[(size_t)i * m + (k - 1)] += h2;
                for (size_t q = 0; q < n; q++) vn[q] -= h2 * vi[q];
            }
            nrm0 -= H[(size_t)i * m + (k - 1)] * H[(size_t)i * m + (k - 1)];
            if (nrm0 < 0.0) nrm0 = 0.0;
            thr = nrm0 * 0.99 * 0.99;
        }
        double hn = sqrt(dot(vn, vn, n));
        H[(size_t)k * m + (k - 1)] = hn;
        if (hn > 0.0) for (size_t q = 0; q < n; q++) vn[q] *= 1.0 / hn;
        /* apply previous rotations, generate a new one */
        for (int i = 0; i < k - 1; i++) {
            double a = H[(size_t)i * m + (k - 1)], b = H[(size_t)(i + 1) * m + (k - 1)];
            H[(size_t)i * m + (k - 1)] = cs[i] * a + sn[i] * b;
            H[(size_t)(i + 1) * m + (k - 1)] = cs[i] * b - sn[i] * a;
        }
        double a = H[(size_t)(k - 1) * m + (k - 1)], b = H[(size_t)k * m + (k - 1)];
        double rr = sqrt(a * a + b * b), c = 1.0, s = 0.0;
        if (rr > 0.0) { c = a / rr; s = b / rr; }
        cs[k - 1] = c; sn[k - 1] = s;
        H[(size_t)(k - 1) * m + (k - 1)] = rr;
        gv[k] = -s * gv[k - 1];
        gv[k - 1] = c * gv[k - 1];
        est = fabs(gv[k]);
        *res = est;
        if (!(k < m && !(est <= tol) && matvecs < m)) break;
        if (hn == 0.0) break;
    }
    /* back substitution, update */
    {
        double *y = xcalloc(k + 1, sizeof(double));
        int kk = k;
        while (kk > 0 && H[(size_t)(kk - 1) * m + (kk - 1)] == 0.0) kk--;
        for (int i = kk - 1; i >= 0; i--) {
            double s = gv[i];
            for (int q = i + 1; q < kk; q++) s -= H[(size_t)i * m + q] * y[q];
            y[i] = s / H[(size_t)i * m + i];
        }
        const double *B = flex ? Z : V;
        for (int i = 0; i < kk; i++) { const double *vi = B + (size_t)i * n; for (size_t q = 0; q < n; q++) sol[q] += y[i] * vi[q]; }
        free(y);
    }
    status = (est <= tol) ? 0 : -1;
done:
    *iters = matvecs;
    free(H); free(cs); free(sn); free(gv);
    (void)z;
    return status;
}

/* dc.f:1-151, defect correction: x <- x + M^-1 (b - A x); stop on ||M^-1 r|| <= rtol*||M^-1 r0|| + atol
 * (tested before the correction is added); at most ipar(6) = m matrix-vector products. */
static int dc_solve(owo_tank *t, const double *rhs, double *sol, int *iters, double *res) {
    size_t n = t->g.n;
    double *w = t->v1, *z = t->v2;
    int matvecs = 0; double tol = 0.0;
    while (1) {
        owo_matvec(t, sol, w); matvecs++;
        for (size_t i = 0; i < n; i++) w[i] = rhs[i] - w[i];
        owo_precond(t, w, z);
        double alpha = sqrt(dot(z, z, n));
        if (matvecs == 1) tol = t->p.reltol * alpha + t->p.abstol;
        *res = alpha;
        if (alpha <= tol) { *iters = matvecs; return 0; }
        for (size_t i = 0; i < n; i++) sol[i] = sol[i] + z[i];
        if (matvecs >= t->p.gmres_m) { *iters = matvecs; return -1; }
    }
}

/* iterative_solution.f90:1-102 */
int owo_solve(owo_tank *t, const double *rhs, double *sol, int *iters, double *res) {
    int st = t->p.solver == 0 ? dc_solve(t, rhs, sol, iters, res) : gmres_solve(t, rhs, sol, iters, res);
    t->it_total += *iters; t->it_solves++;
    if (t->it_min < 0 || *iters < t->it_min) t->it_min = *iters;
    if (*iters > t->it_max) t->it_max = *iters;
    return st;
}

void owo_stats(const owo_tank *t, long *out4) {
    out4[0] = t->it_total; out4[1] = t->it_solves; out4[2] = t->it_min; out4[3] = t->it_max;
}

/* ------------------------------------------------------------------------- */
/* RK4 step  (Runge_Kutta_4.f90:1-299)                                        */
/* ------------------------------------------------------------------------- */
static void fill_rhs(owo_tank *t, const double *P) {
    const grid_t *g = &t->g;
    /* Dirichlet data on the free surface of non-ghost columns; zero Neumann data on the
     * x-wall ghost columns (no flux, no paddle); everything else zero (lines 65-69). */
    for (int j = g->gy; j < g->nyg - g->gy; j++)
        for (int i = g->gx; i < g->nxg - g->gx; i++)
            t->rhs[I3(g, g->nzg - 1, i, j)] = P[I2(g, i, j)];
}

int owo_rk4_step(owo_tank *t, double time, double dt) {
    grid_t *g = &t->g;
    size_t nxy = g->nxy, n = g->n;
    int corners = (g->gx + g->gy == 2);
    double *k1E = xcalloc(nxy, 8), *k1P = xcalloc(nxy, 8), *k2E = xcalloc(nxy, 8), *k2P = xcalloc(nxy, 8);
    double *k3E = xcalloc(nxy, 8), *k3P = xcalloc(nxy, 8), *k4E = xcalloc(nxy, 8), *k4P = xcalloc(nxy, 8);
    double *Et = xcalloc(nxy, 8), *Pt = xcalloc(nxy, 8), *Wt = xcalloc(nxy, 8);
    int it, st = 0; double res;
    t->tstep++;
    int idxnew = (int)((t->tstep - 1) % 2), idxlast = (int)(t->tstep % 2);
    memcpy(t->last[idxnew], t->phi, n * sizeof(double));
#define ZC(A, B) do { if (corners) { size_t c4[4] = {I2(g,0,0), I2(g,g->nxg-1,0), I2(g,0,g->nyg-1), I2(g,g->nxg-1,g->nyg-1)}; \
        for (int q = 0; q < 4; q++) { (A)[c4[q]] = 0.0; (B)[c4[q]] = 0.0; } } } while (0)
    /* stage 1 */
    owo_rhs(t, t->E, t->P, t->W, k1E, k1P); ZC(k1E, k1P);
    for (size_t q = 0; q < nxy; q++) { Et[q] = t->E[q] + 0.5 * dt * k1E[q]; Pt[q] = t->P[q] + 0.5 * dt * k1P[q]; }
    relax(t, Et, Pt);
    surface_update(t, Et, Pt);
    fill_rhs(t, Pt);
    if (t->p.extrapolation == 1)
        for (size_t q = 0; q < n; q++) t->phi[q] = 1.5 * t->last[idxnew][q] - t->last[idxlast][q] * 0.5;
    st |= owo_solve(t, t->rhs, t->phi, &it, &res);
    owo_vertical_velocity(t, t->phi, Wt);
    /* stage 2 */
    owo_rhs(t, Et, Pt, Wt, k2E, k2P); ZC(k2E, k2P);
    for (size_t q = 0; q < nxy; q++) { Et[q] = t->E[q] + 0.5 * dt * k2E[q]; Pt[q] = t->P[q] + 0.5 * dt * k2P[q]; }
    relax(t, Et, Pt);
    surface_update(t, Et, Pt);
    fill_rhs(t, Pt);
    st |= owo_solve(t, t->rhs, t->phi, &it, &res);
    owo_vertical_velocity(t, t->phi, Wt);
    /* stage 3 */
    owo_rhs(t, Et, Pt, Wt, k3E, k3P); ZC(k3E, k3P);
    for (size_t q = 0; q < nxy; q++) { Et[q] = t->E[q] + dt * k3E[q]; Pt[q] = t->P[q] + dt * k3P[q]; }
    relax(t, Et, Pt);
    surface_update(t, Et, Pt);
    fill_rhs(t, Pt);
    if (t->p.extrapolation == 1 && t->tstep > 1)
        for (size_t q = 0; q < n; q++)
            t->phi[q] = 2.66667 * t->phi[q] - 2.0 * t->last[idxnew][q] + 0.333333 * t->last[idxlast][q];
    st |= owo_solve(t, t->rhs, t->phi, &it, &res);
    owo_vertical_velocity(t, t->phi, Wt);
    /* stage 4 and update */
    owo_rhs(t, Et, Pt, Wt, k4E, k4P); ZC(k4E, k4P);
    for (size_t q = 0; q < nxy; q++) {
        t->E[q] = t->E[q] + dt / 6.0 * (k1E[q] + 2.0 * k2E[q] + 2.0 * k3E[q] + k4E[q]);
        t->P[q] = t->P[q] + dt / 6.0 * (k1P[q] + 2.0 * k2P[q] + 2.0 * k3P[q] + k4P[q]);
    }
    relax(t, t->E, t->P);
    surface_update(t, t->E, t->P);
    fill_rhs(t, t->P);
    st |= owo_solve(t, t->rhs, t->phi, &it, &res);
    owo_vertical_velocity(t, t->phi, t->W);
#undef ZC
    (void)time;
    free(k1E); free(k1P); free(k2E); free(k2P); free(k3E); free(k3P); free(k4E); free(k4P);
    free(Et); free(Pt); free(Wt);
    return st;
}

/* ------------------------------------------------------------------------- */
/* Construction                                                               */
/* ------------------------------------------------------------------------- */
owo_tank *owo_create(const owo_params *p, const double *h, const double *hx, const double *hy,
                     const double *hxx, const double *hyy) {
    if (p->nx < 1 || p->ny < 1 || p->nz < 3) return NULL;
    if (p->nx == 1) { fprintf(stderr, "ow3d_oracle: y-z problems (Nx == 1) are not supported\n"); return NULL; }
    owo_tank *t = xcalloc(1, sizeof(owo_tank));
    t->p = *p;
    grid_t *g = &t->g;
    g->nx = p->nx; g->ny = p->ny; g->nz = p->nz;
    g->gx = p->nx > 1 ? 1 : 0; g->gy = p->ny > 1 ? 1 : 0;
    g->a = p->nx > 1 ? p->alpha : 0; g->b = p->ny > 1 ? p->beta : 0; g->c = p->gamma;
    if ((p->ny > 1 && (p->alpha != p->beta || p->alpha != p->gamma)) || (p->ny == 1 && p->alpha != p->gamma)) {
        fprintf(stderr, "ow3d_oracle: the mixed-derivative tables need alpha == beta == gamma\n");
        free(t); return NULL;
    }
    if (2 * g->c + 1 - 1 > p->nz) { fprintf(stderr, "ow3d_oracle: sigma stencil wider than the column\n"); free(t); return NULL; }
    grid_alloc(g);
    grid_fine_coords(g, p->lx, p->ly, p->lz, p->gridz);
    size_t nxy = g->nxy, n = g->n;
    memcpy(g->h, h, nxy * sizeof(double));
    if (hx) memcpy(g->hx, hx, nxy * sizeof(double));
    if (hy) memcpy(g->hy, hy, nxy * sizeof(double));
    if (hxx) memcpy(g->hxx, hxx, nxy * sizeof(double));
    if (hyy) memcpy(g->hyy, hyy, nxy * sizeof(double));
    grid_stencils(g);
    t->ds = xcalloc(5 * n, sizeof(double));
    compute_dsigma(g, NULL, NULL, NULL, NULL, NULL, t->ds);
    t->E = xcalloc(nxy, 8); t->P = xcalloc(nxy, 8); t->W = xcalloc(nxy, 8);
    t->Ex = xcalloc(nxy, 8); t->Exx = xcalloc(nxy, 8); t->Ey = xcalloc(nxy, 8); t->Eyy = xcalloc(nxy, 8);
    t->Px = xcalloc(nxy, 8); t->Py = xcalloc(nxy, 8);
    t->phi = xcalloc(n, 8); t->rhs = xcalloc(n, 8); t->last[0] = xcalloc(n, 8); t->last[1] = xcalloc(n, 8);
    for (int q = 0; q < 10; q++) t->tmp[q] = xcalloc(n, 8);
    int m = p->gmres_m > 1 ? p->gmres_m : 15;
    t->kw = xcalloc((size_t)(m + 1) * n, 8);
    if (t->p.gmres_side == 1) t->kz = xcalloc((size_t)m * n, 8);
    t->v1 = xcalloc(n, 8); t->v2 = xcalloc(n, 8); t->v3 = xcalloc(n, 8);
    t->it_min = -1;
    if (p->precond == 3 && mg_setup(t) != 0) { owo_destroy(t); return NULL; }
    return t;
}

void owo_destroy(owo_tank *t) {
    if (!t) return;
    grid_free(&t->g);
    free(t->ds); free(t->E); free(t->P); free(t->W); free(t->Ex); free(t->Exx); free(t->Ey); free(t->Eyy);
    free(t->Px); free(t->Py); free(t->phi); free(t->rhs); free(t->last[0]); free(t->last[1]);
    for (int q = 0; q < 10; q++) free(t->tmp[q]);
    free(t->kw); free(t->kz); free(t->v1); free(t->v2); free(t->v3);
    for (int k = 0; k < t->nlev; k++) {
        level_t *L = &t->lev[k];
        grid_free(&L->g); free(L->ds); free(L->rp); free(L->ci); free(L->va); free(L->ighost);
        free(L->lu); free(L->piv); free(L->x); free(L->b); free(L->r); free(L->t1);
    }
    free(t->lev);
    for (int z = 0; z < t->nzones; z++) free(t->zones[z].gam);
    free(t->zones);
    free(t);
}

int owo_dims(const owo_tank *t, int *o) {
    o[0] = t->g.nxg; o[1] = t->g.nyg; o[2] = t->g.nzg; o[3] = t->nlev; o[4] = (int)t->g.n; o[5] = t->g.nxy;
    return 0;
}
void owo_grid(const owo_tank *t, double *x, double *y, double *z) {
    memcpy(x, t->g.x, t->g.nxg * 8); memcpy(y, t->g.y, t->g.nyg * 8); memcpy(z, t->g.z, t->g.nzg * 8);
}
void owo_stencils(const owo_tank *t, double *sx, double *sy, double *sz) {
    const grid_t *g = &t->g;
    if (sx) memcpy(sx, g->sx, (size_t)g->nxg * 2 * g->ra * 8);
    if (sy) memcpy(sy, g->sy, (size_t)g->nyg * 2 * g->rb * 8);
    if (sz) memcpy(sz, g->sz, (size_t)g->nzg * 2 * g->rc * 8);
}

/* Initial state.  As in OceanWave3DT0Setup.f90:444, the plane derivatives (and, for nonlinear
 * runs, the sigma coefficients) are set up from the initial surface before the first step. */
void owo_set_state(owo_tank *t, const double *E, const double *P, const double *W, const double *phi) {
    size_t nxy = t->g.nxy, n = t->g.n;
    memcpy(t->E, E, nxy * 8); memcpy(t->P, P, nxy * 8);
    if (W) memcpy(t->W, W, nxy * 8); else memset(t->W, 0, nxy * 8);
    if (phi) memcpy(t->phi, phi, n * 8); else memset(t->phi, 0, n * 8);
    surface_update(t, t->E, t->P);
}
void owo_get_state(const owo_tank *t, double *E, double *P, double *W, double *phi) {
    size_t nxy = t->g.nxy, n = t->g.n;
    if (E) memcpy(E, t->E, nxy * 8);
    if (P) memcpy(P, t->P, nxy * 8);
    if (W) memcpy(W, t->W, nxy * 8);
    if (phi) memcpy(phi, t->phi, n * 8);
}
