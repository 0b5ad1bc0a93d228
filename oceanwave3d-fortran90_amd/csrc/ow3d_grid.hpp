// Host-side set-up of grids, finite-difference stencils and the multigrid hierarchy.
// Follows the reference's set-up for the path:
//   SetupCompDomain.f90:35-79 (coordinates), BuildStencilsGridX/Y.f90 and
//   PreProcessDiffStencils.f90:57-86 (stencils via Fornberg's recursion, as the reference's
//   `weights`, utilities/stream_func_wave_finite.f:128-168), MGPreProcess.f90:30-318
//   (levels, coarsening strategy 0, coarse coordinates and depths).
#pragma once
#include <cmath>
#include <cstddef>
#include <string>
#include <vector>

namespace ow3d {

struct Grid {
    int nx = 0, ny = 0, nz = 0;     // interior counts (deck convention)
    int gx = 0, gy = 0;             // wall ghost layers in x / y
    int nxg = 0, nyg = 0, nzg = 0, nxy = 0;
    size_t n = 0;
    int a = 0, b = 0, c = 0;        // half widths (a = 0 if nx == 1, b = 0 if ny == 1)
    int ra = 1, rb = 1, rc = 1;
    int xc = 0, yc = 0, zc = 0;     // coarsened from the next finer level in x / y / z
    std::vector<double> x, y, z, h, hx, hy, hxx, hyy;
    std::vector<double> sx, sy, sz; // [N][2][r]: first then second derivative rows

    void alloc() {
        nxg = nx + 2 * gx; nyg = ny + 2 * gy; nzg = nz + 1; nxy = nxg * nyg;
        n = (size_t)nxy * nzg;
        ra = 2 * a + 1; rb = 2 * b + 1; rc = 2 * c + 1;
        x.assign(nxg, 0.0); y.assign(nyg, 0.0); z.assign(nzg, 0.0);
        h.assign(nxy, 0.0); hx.assign(nxy, 0.0); hy.assign(nxy, 0.0);
        hxx.assign(nxy, 0.0); hyy.assign(nxy, 0.0);
        sx.assign((size_t)nxg * 2 * ra, 0.0); sy.assign((size_t)nyg * 2 * rb, 0.0);
        sz.assign((size_t)nzg * 2 * rc, 0.0);
    }
};

// Fornberg (Math. Comp. 51, 1988) weights for derivatives 0..m at z on nodes x[0..n].
inline void fornberg(double z, const double *x, int n, int m, double *c) {
    auto C = [&](int j, int k) -> double & { return c[j * (m + 1) + k]; };
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
                for (int k = mn; k >= 1; k--) C(i, k) = c1 * ((double)k * C(i - 1, k - 1) - c5 * C(i - 1, k)) / c2;
                C(i, 0) = -(c1 * c5 * C(i - 1, 0) / c2);
            }
            for (int k = mn; k >= 1; k--) C(j, k) = (c4 * C(j, k) - (double)k * C(j, k - 1)) / c3;
            C(j, 0) = c4 * C(j, 0) / c3;
        }
        c1 = c2;
    }
}

inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// One stencil row per node; the window of node i is [clamp(i-half,0,N-r), +r).
inline void build_stencils(const std::vector<double> &x, int N, int half, std::vector<double> &S) {
    int r = 2 * half + 1;
    std::vector<double> c((size_t)r * 3);
    for (int i = 0; i < N; i++) {
        int ws = clampi(i - half, 0, N - r);
        for (int o = 1; o <= 2; o++) {
            fornberg(x[i], x.data() + ws, r - 1, o, c.data());
            for (int m = 0; m < r; m++) S[(size_t)(i * 2 + (o - 1)) * r + m] = c[(size_t)m * (o + 1) + o];
        }
    }
}

inline void grid_stencils(Grid &g) {
    if (g.nx > 1) build_stencils(g.x, g.nxg, g.a, g.sx);
    if (g.ny > 1) build_stencils(g.y, g.nyg, g.b, g.sy);
    build_stencils(g.z, g.nzg, g.c, g.sz);
}

inline void fine_coords(Grid &g, double lx, double ly, double lz, int gridz) {
    if (g.nx == 1) g.x[0] = lx;
    else for (int ii = 0; ii < g.nxg; ii++) g.x[ii] = ((double)(ii - g.gx) / (double)(g.nx - 1)) * lx;
    if (g.ny == 1) g.y[0] = ly;
    else for (int jj = 0; jj < g.nyg; jj++) g.y[jj] = ((double)(jj - g.gy) / (double)(g.ny - 1)) * ly;
    for (int k = 0; k < g.nz; k++) {
        if (gridz == 1) g.z[k + 1] = std::sin(M_PI * (double)k / (double)(2 * (g.nz - 1)));
        else g.z[k + 1] = (double)k / (double)(g.nz - 1) * lz;
    }
    g.z[0] = -g.z[2];
}

// Number of multigrid levels (MGPreProcess.f90:449-458; evaluated in single precision there).
inline int no_grids(int nx, int ny, int nz) {
    int m = nx > ny ? nx : ny; m = m > nz ? m : nz;
    float tmp = (float)(std::log10((float)(m - 1)) / std::log10(2.0));
    return (int)std::lround(tmp);
}

// Directions halved from an (nx, ny, nz) level: coarsening strategy 0 of MGPreProcess.f90:140-178 (halve
// x / y while larger than Nz, all three when equal).  Where that would halve an even column count (no
// nested coarse sigma grid; the reference stops with an error, MGPreProcess.f90:183-211) the column is
// kept and x / y are halved while they have an odd count above 3.  Returns false when nothing can be halved.
inline bool coarsen_flags(int nx, int ny, int nz, int &xc, int &yc, int &zc) {
    xc = yc = zc = 0;
    if (nx == 1) { if (ny == nz) { yc = 1; zc = 1; } else yc = 1; }
    else if (ny == 1) { if (nx == nz) { xc = 1; zc = 1; } else xc = 1; }
    else if (nx == ny && nx == nz) { xc = yc = zc = 1; }
    else { if (nx > nz) xc = 1; if (ny > nz) yc = 1; if (nx < nz && ny < nz) zc = 1; }
    if (zc && nz % 2 == 0) {
        zc = 0;
        xc = (nx > 3 && nx % 2 == 1) ? 1 : 0;
        yc = (ny > 3 && ny % 2 == 1) ? 1 : 0;
    }
    return xc || yc || zc;
}

// Builds levels[0] (coarsest) .. levels[L-1] (finest, = fine grid with 2nd-order stencils).
inline bool mg_levels(const Grid &fine, int maxgrids, std::vector<Grid> &lev, std::string &err) {
    int L = no_grids(fine.nx, fine.ny, fine.nz);
    if (maxgrids > 0 && L > maxgrids) L = maxgrids;
    if (L < 1) L = 1;
    // coarsening strategy 0 with the even-column rule of coarsen_flags: stop when nothing can be halved
    {
        int nx = fine.nx, ny = fine.ny, nz = fine.nz;
        for (int q = 1; q < L; q++) {
            int xc, yc, zc;
            if (!coarsen_flags(nx, ny, nz, xc, yc, zc)) { L = q; break; }
            if (xc) nx = (nx - 1) / 2 + 1;
            if (yc) ny = (ny - 1) / 2 + 1;
            if (zc) nz = (nz - 1) / 2 + 1;
        }
    }
    lev.assign(L, Grid());
    for (int k = L - 1; k >= 0; k--) {
        Grid &g = lev[k];
        if (k == L - 1) { g.nx = fine.nx; g.ny = fine.ny; g.nz = fine.nz; }
        else {
            const Grid &f = lev[k + 1];
            int NFx = f.nx, NFy = f.ny, NFz = f.nz, xc = 0, yc = 0, zc = 0;
            coarsen_flags(NFx, NFy, NFz, xc, yc, zc);
            if (xc && NFx % 2 == 0) { err = "multigrid needs odd Nx on every coarsened level"; return false; }
            if (yc && NFy % 2 == 0) { err = "multigrid needs odd Ny on every coarsened level"; return false; }
            if (zc && NFz % 2 == 0) { err = "multigrid needs odd Nz on every coarsened level"; return false; }
            g.nx = xc ? (NFx - 1) / 2 + 1 : NFx;
            g.ny = yc ? (NFy - 1) / 2 + 1 : NFy;
            g.nz = zc ? (NFz - 1) / 2 + 1 : NFz;
            g.xc = xc; g.yc = yc; g.zc = zc;
        }
        g.gx = fine.gx; g.gy = fine.gy;
        g.a = g.nx > 1 ? 1 : 0; g.b = g.ny > 1 ? 1 : 0; g.c = 1;
        g.alloc();
        if (k == L - 1) {
            g.x = fine.x; g.y = fine.y; g.z = fine.z;
            g.h = fine.h; g.hx = fine.hx; g.hy = fine.hy; g.hxx = fine.hxx; g.hyy = fine.hyy;
        } else {
            const Grid &f = lev[k + 1];
            int XF = g.xc ? 2 : 1, YF = g.yc ? 2 : 1, ZF = g.zc ? 2 : 1;
            if (g.nx > 1) {
                for (int ic = 0; ic < g.nx; ic++) g.x[g.gx + ic] = f.x[f.gx + ic * XF];
                if (g.gx) {
                    g.x[0] = g.x[1] - (g.x[2] - g.x[1]);
                    g.x[g.nxg - 1] = g.x[g.nxg - 2] + (g.x[g.nxg - 2] - g.x[g.nxg - 3]);
                }
            } else g.x[0] = f.x[0];
            if (g.ny > 1) {
                for (int jc = 0; jc < g.ny; jc++) g.y[g.gy + jc] = f.y[f.gy + jc * YF];
                if (g.gy) {
                    g.y[0] = g.y[1] - (g.y[2] - g.y[1]);
                    g.y[g.nyg - 1] = g.y[g.nyg - 2] + (g.y[g.nyg - 2] - g.y[g.nyg - 3]);
                }
            } else g.y[0] = f.y[0];
            g.z[0] = -f.z[1 + ZF];
            for (int kc = 0; kc < g.nz; kc++) g.z[1 + kc] = f.z[1 + kc * ZF];
            for (int jc = 0; jc < g.ny; jc++)
                for (int ic = 0; ic < g.nx; ic++) {
                    size_t pc = (size_t)(g.gx + ic) + (size_t)g.nxg * (g.gy + jc);
                    size_t pf = (size_t)(f.gx + ic * XF) + (size_t)f.nxg * (f.gy + jc * YF);
                    g.h[pc] = f.h[pf]; g.hx[pc] = f.hx[pf]; g.hxx[pc] = f.hxx[pf];
                    g.hy[pc] = f.hy[pf]; g.hyy[pc] = f.hyy[pf];
                }
            for (int j = 0; j < g.nyg; j++) {
                g.h[(size_t)g.nxg * j] = g.h[(size_t)g.gx + (size_t)g.nxg * j];
                g.h[(size_t)(g.nxg - 1) + (size_t)g.nxg * j] = g.h[(size_t)(g.nxg - 1 - g.gx) + (size_t)g.nxg * j];
            }
            for (int i = 0; i < g.nxg; i++) {
                g.h[i] = g.h[(size_t)i + (size_t)g.nxg * g.gy];
                g.h[(size_t)i + (size_t)g.nxg * (g.nyg - 1)] = g.h[(size_t)i + (size_t)g.nxg * (g.nyg - 1 - g.gy)];
            }
        }
        grid_stencils(g);
    }
    return true;
}

}  // namespace ow3d
