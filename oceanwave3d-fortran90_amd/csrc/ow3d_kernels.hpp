// HIP kernels for gfx950 (MI355X) of the Laplace-solve-per-RK4-stage path.
//
// Layout (as the reference, BuildLinearSystem.f90:117): sigma-fastest columns,
//   q = k + nzg*(i + nxg*j);  plane fields p = i + nxg*j.
// The high-order matvec keeps whole sigma columns of a 2D tile (plus an alpha+1 halo) in LDS;
// every interior column's x-, y-, sigma- and mixed derivatives are formed from that tile, so
// phi is read from HBM once per matvec (halo re-reads hit L2 / MALL) and y written once.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace ow3d {

__host__ __device__ constexpr int cmin(int a, int b) { return a < b ? a : b; }
__host__ __device__ constexpr int cmax(int a, int b) { return a > b ? a : b; }
__host__ __device__ constexpr int cclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// ---------------------------------------------------------------------------------------------
// Sigma-transformation coefficients at one point (DetermineTransformationConstantsArray.f90:143-205)
// ---------------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void sigma_coef3(T s1, T d, T hx, T hy, T hs, T Ex, T Ey, T Es, T &s2, T &s3, T &s4, T &s5) {
    const T one = T(1), two = T(2);
    s2 = ((one - s1) * hx - s1 * Ex) / d;
    s3 = -(two * (one - s1) * (hx * hx + hy * hy) / (d * d))
         + two * (two * s1 - one) * (Ex * hx + Ey * hy) / (d * d)
         + (one - s1) * hs / d
         + two * s1 * (Ex * Ex + Ey * Ey) / (d * d)
         - s1 * Es / d;
    s4 = ((one - s1) * hy - s1 * Ey) / d;
    s5 = one / d;
}
template <typename T>
__device__ __forceinline__ void sigma_coef2(T s1, T d, T hx, T hs, T Ex, T Es, T &s2, T &s3, T &s5) {
    const T one = T(1), two = T(2);
    s2 = ((one - s1) * hx - s1 * Ex) / d;
    s3 = -(two * (one - s1) * (hx * hx) / (d * d))
         + two * (two * s1 - one) * (Ex * hx) / (d * d)
         + (one - s1) * hs / d
         + two * s1 * (Ex * Ex) / (d * d)
         - s1 * Es / d;
    s5 = one / d;
}

// ---------------------------------------------------------------------------------------------
// High-order matvec  (BuildLinearSystem.f90:1-290)
// ---------------------------------------------------------------------------------------------
struct MvArgs {
    int nxg, nyg, nxy;
    const double *sx, *sy, *sz;  // stencil rows [N][2][2A+1]
    const double *zs;            // sigma of each level (0 on the ghost level)
    const double *cf;            // 7 planes: sigma coefficients affine in sigma, s2 = A2 + sigma*B2, s4 = A4 + sigma*B4,
                                 // s3 = A3 + sigma*B3, s5 = S5; planes (A2, B2, A4, B4, A3, B3, S5)
};

// Per-column coefficient planes from the surface and the bottom: DetermineTransformationConstantsArray.f90
// (lines 190-201) written as functions affine in sigma, evaluated once per column and stage.
__device__ __forceinline__ void column_coefs(double d, double Ex, double Ey, double Es, double hx, double hy, double hs,
                                             double *cf, int nxy, int p) {
    const double id = 1.0 / d, id2 = id * id;
    const double H2 = hx * hx + hy * hy, C = Ex * hx + Ey * hy, E2 = Ex * Ex + Ey * Ey;
    cf[p] = hx * id;
    cf[nxy + p] = -(hx + Ex) * id;
    cf[2 * nxy + p] = hy * id;
    cf[3 * nxy + p] = -(hy + Ey) * id;
    cf[4 * nxy + p] = (-2.0 * H2 - 2.0 * C) * id2 + hs * id;
    cf[5 * nxy + p] = (2.0 * H2 + 4.0 * C + 2.0 * E2) * id2 - (hs + Es) * id;
    cf[6 * nxy + p] = id;
}

// 3D, one thread per grid point: block = 16 x TY columns x NZ levels, NT threads.  The tile is
// stored k-fastest in LDS exactly as in HBM (row segments are straight copies), and threads are mapped
// k-fastest too, so every LDS access of a wave and every global load / store is contiguous.
// Phase 1 forms the x- and y-derivatives of every tile point (the first passes of the mixed
// derivatives go to LDS); phase 2 the sigma derivatives, the mixed derivatives and the row.
template <int NZ, int A, int TY, int NT>
__global__ __launch_bounds__(NT) void k_matvec3p(MvArgs g, const double *__restrict__ phi, double *__restrict__ out) {
    constexpr int TX = 16, RA = 2 * A + 1, R = A + 1, WX = TX + 2 * R, WY = TY + 2 * R, NC = TX * TY, NP = NC * NZ;
    constexpr int PER = (NP + NT - 1) / NT;
    __shared__ double sP[WX * WY * NZ];
    __shared__ double sDx[NP];
    __shared__ double sDy[NP];
    __shared__ double sSX[(TX + 2) * 2 * RA];
    __shared__ double sSY[(TY + 2) * 2 * RA];
    __shared__ double sSZ[NZ * 2 * RA];
    __shared__ double sZS[NZ];
    const int nxg = g.nxg, nyg = g.nyg, nxy = g.nxy;
    const int i0 = blockIdx.x * TX, j0 = blockIdx.y * TY;
    const int lox = cmax(cmin(i0 - R, nxg - 2 - 2 * A), 0);
    const int loy = cmax(cmin(j0 - R, nyg - 2 - 2 * A), 0);
    const int wx = cmin(WX, nxg - lox), wy = cmin(WY, nyg - loy);
    for (int jj = 0; jj < wy; jj++) {
        const double *src = phi + ((size_t)(loy + jj) * nxg + lox) * NZ;
        double *dst = sP + jj * WX * NZ;
        for (int e = threadIdx.x; e < wx * NZ; e += NT) dst[e] = src[e];
    }
    for (int e = threadIdx.x; e < (TX + 2) * 2 * RA; e += NT) {
        const int r = e / (2 * RA), ii = i0 - 1 + r;
        if (ii >= 0 && ii < nxg) sSX[e] = g.sx[(size_t)ii * 2 * RA + (e - r * 2 * RA)];
    }
    for (int e = threadIdx.x; e < (TY + 2) * 2 * RA; e += NT) {
        const int r = e / (2 * RA), jj = j0 - 1 + r;
        if (jj >= 0 && jj < nyg) sSY[e] = g.sy[(size_t)jj * 2 * RA + (e - r * 2 * RA)];
    }
    for (int e = threadIdx.x; e < NZ * 2 * RA; e += NT) sSZ[e] = g.sz[e];
    if (threadIdx.x < NZ) sZS[threadIdx.x] = g.zs[threadIdx.x];
    __syncthreads();
    auto P = [&](int k, int ii, int jj) -> double { return sP[((jj - loy) * WX + (ii - lox)) * NZ + k]; };
    auto SX = [&](int ii) -> const double * { return sSX + (ii - (i0 - 1)) * 2 * RA; };
    auto SY = [&](int jj) -> const double * { return sSY + (jj - (j0 - 1)) * 2 * RA; };
    double acc[PER];
#pragma unroll
    for (int r = 0; r < PER; r++) {
        acc[r] = 0.0;
        const int pt = threadIdx.x + r * NT;
        if (pt >= NP) break;
        const int col = pt / NZ, k = pt - col * NZ;
        const int i = i0 + col % TX, j = j0 + col / TX;
        if (i >= nxg || j >= nyg || i == 0 || i == nxg - 1 || j == 0 || j == nyg - 1) continue;
        const int wsx = cclamp(i - A, 0, nxg - RA), wsy = cclamp(j - A, 0, nyg - RA);
        double d1 = 0.0, d2 = 0.0, e1 = 0.0, e2 = 0.0;
        {
            const double *S1 = SX(i), *S2 = S1 + RA;
#pragma unroll
            for (int m = 0; m < RA; m++) { const double v = P(k, wsx + m, j); d1 += S1[m] * v; d2 += S2[m] * v; }
        }
        {
            const double *S1 = SY(j), *S2 = S1 + RA;
#pragma unroll
            for (int m = 0; m < RA; m++) { const double v = P(k, i, wsy + m); e1 += S1[m] * v; e2 += S2[m] * v; }
        }
        if (k == 0) {  // corner-corrected first pass on the sigma ghost level (ConstructTableCrossDerivatives.f90:173-186)
            int sh = 0;
            if (wsx == 0) sh = 1;
            if (wsx + RA - 1 == nxg - 1) sh = -1;
            if (sh != 0) {
                const double *S = SX(i - sh);
                double t = 0.0;
#pragma unroll
                for (int m = 0; m < RA; m++) t += S[m] * P(0, wsx + sh + m, j);
                d1 = t;
            }
            sh = 0;
            if (wsy == 0) sh = 1;
            if (wsy + RA - 1 == nyg - 1) sh = -1;
            if (sh != 0) {
                const double *S = SY(j - sh);
                double t = 0.0;
#pragma unroll
                for (int m = 0; m < RA; m++) t += S[m] * P(0, i, wsy + sh + m);
                e1 = t;
            }
        }
        sDx[pt] = d1;
        sDy[pt] = e1;
        acc[r] = d2 + e2;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < PER; r++) {
        const int pt = threadIdx.x + r * NT;
        if (pt >= NP) break;
        const int col = pt / NZ, k = pt - col * NZ;
        const int i = i0 + col % TX, j = j0 + col / TX;
        if (i >= nxg || j >= nyg) continue;
        const bool wxl = (i == 0 || i == nxg - 1), wyl = (j == 0 || j == nyg - 1);
        double o;
        if ((wxl && wyl) || ((wxl || wyl) && k == 0) || (!wxl && !wyl && k == NZ - 1)) {
            o = P(k, i, j);
        } else if (wxl) {
            const int ib = (i == 0) ? 1 : nxg - 2, ws = cclamp(ib - A, 0, nxg - RA);
            const double *S = SX(ib);
            o = 0.0;
#pragma unroll
            for (int m = 0; m < RA; m++) o += S[m] * P(k, ws + m, j);
        } else if (wyl) {
            const int jb = (j == 0) ? 1 : nyg - 2, ws = cclamp(jb - A, 0, nyg - RA);
            const double *S = SY(jb);
            o = 0.0;
#pragma unroll
            for (int m = 0; m < RA; m++) o += S[m] * P(k, i, ws + m);
        } else {
            const int p = i + nxg * j;
            const double A2 = g.cf[p], B2 = g.cf[nxy + p], A4 = g.cf[2 * nxy + p], B4 = g.cf[3 * nxy + p];
            const double S5 = g.cf[6 * nxy + p];
            const double *dxc = sDx + col * NZ, *dyc = sDy + col * NZ;
            if (k == 0) {
                const int wz = cclamp(1 - A, 0, NZ - RA);
                const double *Z1 = sSZ + (1 * 2) * RA;
                double dz = 0.0;
#pragma unroll
                for (int m = 0; m < RA; m++) dz += Z1[m] * P(wz + m, i, j);
                // h_x = A2 * d and h_y = A4 * d with d = 1 / S5: no bottom-slope planes are read
                const double sg = sZS[1], d = 1.0 / S5, hx = A2 * d, hy = A4 * d;
                o = (S5 + hx * (A2 + sg * B2) + hy * (A4 + sg * B4)) * dz + hx * dxc[1] + hy * dyc[1];
            } else {
                const double A3 = g.cf[4 * nxy + p], B3 = g.cf[5 * nxy + p];
                const int wz = cclamp(k - A, 0, NZ - RA);
                const double *Z1 = sSZ + (k * 2) * RA, *Z2 = Z1 + RA;
                double dz = 0.0, dzz = 0.0, dxz = 0.0, dyz = 0.0;
#pragma unroll
                for (int m = 0; m < RA; m++) {
                    const double z1 = Z1[m], cv = P(wz + m, i, j);
                    dz += z1 * cv;
                    dzz += Z2[m] * cv;
                    dxz += z1 * dxc[wz + m];
                    dyz += z1 * dyc[wz + m];
                }
                const double sg = sZS[k];
                const double s2 = A2 + sg * B2, s4 = A4 + sg * B4, s3 = A3 + sg * B3;
                o = acc[r] + s3 * dz + 2.0 * (s2 * dxz + s4 * dyz) + (s2 * s2 + s4 * s4 + S5 * S5) * dzz;
            }
        }
        out[((size_t)j * nxg + i) * NZ + k] = o;
    }
}

// x-z problems (Ny == 1), one thread per grid point: block = 64 columns x NZ levels.
template <int NZ, int A>
__global__ __launch_bounds__(256) void k_matvec2p(MvArgs g, const double *__restrict__ phi, double *__restrict__ out) {
    constexpr int NT = 256, TX = 64, RA = 2 * A + 1, R = A + 1, WX = TX + 2 * R, NP = TX * NZ, PER = (NP + NT - 1) / NT;
    __shared__ double sP[WX * NZ];
    __shared__ double sDx[NP];
    const int nxg = g.nxg, nxy = g.nxy;
    const int i0 = blockIdx.x * TX;
    const int lox = cmax(cmin(i0 - R, nxg - 2 - 2 * A), 0);
    const int wx = cmin(WX, nxg - lox);
    for (int e = threadIdx.x; e < wx * NZ; e += NT) sP[e] = phi[(size_t)lox * NZ + e];
    __syncthreads();
    auto P = [&](int k, int ii) -> double { return sP[(ii - lox) * NZ + k]; };
    double acc[PER];
#pragma unroll
    for (int r = 0; r < PER; r++) {
        acc[r] = 0.0;
        const int pt = threadIdx.x + r * NT;
        if (pt >= NP) break;
        const int col = pt / NZ, k = pt - col * NZ;
        const int i = i0 + col;
        if (i >= nxg || i == 0 || i == nxg - 1) continue;
        const int wsx = cclamp(i - A, 0, nxg - RA);
        const double *S1 = g.sx + (size_t)(i * 2) * RA, *S2 = S1 + RA;
        double d1 = 0.0, d2 = 0.0;
#pragma unroll
        for (int m = 0; m < RA; m++) { const double v = P(k, wsx + m); d1 += S1[m] * v; d2 += S2[m] * v; }
        if (k == 0) {
            int sh = 0;
            if (wsx == 0) sh = 1;
            if (wsx + RA - 1 == nxg - 1) sh = -1;
            if (sh != 0) {
                const double *S = g.sx + (size_t)((i - sh) * 2) * RA;
                double t = 0.0;
#pragma unroll
                for (int m = 0; m < RA; m++) t += S[m] * P(0, wsx + sh + m);
                d1 = t;
            }
        }
        sDx[pt] = d1;
        acc[r] = d2;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < PER; r++) {
        const int pt = threadIdx.x + r * NT;
        if (pt >= NP) break;
        const int col = pt / NZ, k = pt - col * NZ;
        const int i = i0 + col;
        if (i >= nxg) continue;
        double o;
        if (i == 0 || i == nxg - 1) {
            if (k == 0) o = P(0, i);
            else {
                const int ib = (i == 0) ? 1 : nxg - 2, ws = cclamp(ib - A, 0, nxg - RA);
                const double *S = g.sx + (size_t)(ib * 2) * RA;
                o = 0.0;
#pragma unroll
                for (int m = 0; m < RA; m++) o += S[m] * P(k, ws + m);
            }
        } else if (k == NZ - 1) {
            o = P(k, i);
        } else {
            const int p = i;
            const double A2 = g.cf[p], B2 = g.cf[nxy + p], S5 = g.cf[6 * nxy + p];
            const double *dxc = sDx + col * NZ;
            if (k == 0) {
                const int wz = cclamp(1 - A, 0, NZ - RA);
                const double *Z1 = g.sz + (size_t)(1 * 2) * RA;
                double dz = 0.0;
#pragma unroll
                for (int m = 0; m < RA; m++) dz += Z1[m] * P(wz + m, i);
                const double hx = A2 * (1.0 / S5), sg = g.zs[1];
                o = hx * dxc[1] + (S5 + hx * (A2 + sg * B2)) * dz;
            } else {
                const double A3 = g.cf[4 * nxy + p], B3 = g.cf[5 * nxy + p];
                const int wz = cclamp(k - A, 0, NZ - RA);
                const double *Z1 = g.sz + (size_t)(k * 2) * RA, *Z2 = Z1 + RA;
                double dz = 0.0, dzz = 0.0, dxz = 0.0;
#pragma unroll
                for (int m = 0; m < RA; m++) {
                    const double z1 = Z1[m], cv = P(wz + m, i);
                    dz += z1 * cv;
                    dzz += Z2[m] * cv;
                    dxz += z1 * dxc[wz + m];
                }
                const double sg = g.zs[k];
                const double s2 = A2 + sg * B2, s3 = A3 + sg * B3;
                o = acc[r] + 2.0 * s2 * dxz + (s2 * s2 + S5 * S5) * dzz + s3 * dz;
            }
        }
        out[(size_t)i * NZ + k] = o;
    }
}

// ---------------------------------------------------------------------------------------------
// Free-surface plane: ghost update + derivatives  (DifferentiationsFreeSurfacePlane.f90:12-32,
// UpdateGhostLayerX.f90, UpdateGhostLayerY.f90), sigma-coefficient fields for the matvec.
// ---------------------------------------------------------------------------------------------
struct PlaneArgs {
    int nxg, nyg, nxy, gx, gy, a, b, nonlinear;
    const double *sx, *sy, *h;
    const double *bath;   // 3 planes: h_x, h_y, h_xx+h_yy
    double *E, *P, *Ex, *Exx, *Ey, *Eyy, *Px, *Py, *cf;
};

// one block per row j: x ghost values of E and P (rows with j in [gy, nyg-gy)), then x derivatives.
__global__ __launch_bounds__(256) void k_surface_x(PlaneArgs g) {
    const int j = blockIdx.x, nxg = g.nxg, ra = 2 * g.a + 1;
    double *E = g.E + (size_t)j * nxg, *P = g.P + (size_t)j * nxg;
    if (j >= g.gy && j < g.nyg - g.gy && threadIdx.x < 4) {
        double *v = (threadIdx.x & 1) ? P : E;
        if (threadIdx.x < 2) {
            const double *st = g.sx + (size_t)(1 * 2) * ra;
            double s = 0.0;
            for (int m = 1; m < ra; m++) s += st[m] * v[m];
            v[0] = (0.0 - s) / st[0];
        } else {
            const double *st = g.sx + (size_t)((nxg - 2) * 2) * ra;
            double s = 0.0;
            for (int m = 0; m < ra - 1; m++) s += st[m] * v[nxg - ra + m];
            v[nxg - 1] = (0.0 - s) / st[ra - 1];
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nxg; i += blockDim.x) {
        const int ws = cclamp(i - g.a, 0, nxg - ra);
        const double *S1 = g.sx + (size_t)(i * 2) * ra, *S2 = S1 + ra;
        double e1 = 0.0, e2 = 0.0, p1 = 0.0;
        for (int m = 0; m < ra; m++) { e1 += S1[m] * E[ws + m]; e2 += S2[m] * E[ws + m]; p1 += S1[m] * P[ws + m]; }
        const size_t q = (size_t)j * nxg + i;
        g.Ex[q] = e1; g.Exx[q] = e2; g.Px[q] = p1;
        if (g.nyg == 1) {  // x-z problem: coefficient fields complete here
            const double d = g.nonlinear ? E[i] + g.h[q] : g.h[q];
            column_coefs(d, g.nonlinear ? e1 : 0.0, 0.0, g.nonlinear ? e2 : 0.0, g.bath[q], 0.0, g.bath[2 * g.nxy + q], g.cf, g.nxy, (int)q);
        }
    }
}

// one block per column i: y ghost values (columns with i in [gx, nxg-gx)), y derivatives, coefficients.
__global__ __launch_bounds__(256) void k_surface_y(PlaneArgs g) {
    const int i = blockIdx.x, nxg = g.nxg, nyg = g.nyg, rb = 2 * g.b + 1;
    if (i >= g.gx && i < nxg - g.gx && threadIdx.x < 4) {
        double *v = (threadIdx.x & 1) ? g.P : g.E;
        if (threadIdx.x < 2) {
            const double *st = g.sy + (size_t)(1 * 2) * rb;
            double s = 0.0;
            for (int m = 1; m < rb; m++) s += st[m] * v[i + (size_t)nxg * m];
            v[i] = (0.0 - s) / st[0];
        } else {
            const double *st = g.sy + (size_t)((nyg - 2) * 2) * rb;
            double s = 0.0;
            for (int m = 0; m < rb - 1; m++) s += st[m] * v[i + (size_t)nxg * (nyg - rb + m)];
            v[i + (size_t)nxg * (nyg - 1)] = (0.0 - s) / st[rb - 1];
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < nyg; j += blockDim.x) {
        const int ws = cclamp(j - g.b, 0, nyg - rb);
        const double *S1 = g.sy + (size_t)(j * 2) * rb, *S2 = S1 + rb;
        double e1 = 0.0, e2 = 0.0, p1 = 0.0;
        for (int m = 0; m < rb; m++) {
            const size_t qq = i + (size_t)nxg * (ws + m);
            e1 += S1[m] * g.E[qq]; e2 += S2[m] * g.E[qq]; p1 += S1[m] * g.P[qq];
        }
        const size_t q = i + (size_t)nxg * j;
        g.Ey[q] = e1; g.Eyy[q] = e2; g.Py[q] = p1;
        if (g.nonlinear)
            column_coefs(g.E[q] + g.h[q], g.Ex[q], e1, g.Exx[q] + e2, g.bath[q], g.bath[g.nxy + q], g.bath[2 * g.nxy + q], g.cf, g.nxy, (int)q);
        else
            column_coefs(g.h[q], 0.0, 0.0, 0.0, g.bath[q], g.bath[g.nxy + q], g.bath[2 * g.nxy + q], g.cf, g.nxy, (int)q);
    }
}

// W = dsigma/dz * one-sided sigma derivative at the surface (VerticalFreeSurfaceVelocity.f90:17).
__global__ void k_vertical_velocity(int nxy, int nzg, int rc, const double *__restrict__ szTop,
                                    const double *__restrict__ cf, const double *__restrict__ phi, double *__restrict__ W) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nxy) return;
    const double *col = phi + (size_t)p * nzg + (nzg - rc);
    double s = 0.0;
    for (int m = 0; m < rc; m++) s += szTop[m] * col[m];
    W[p] = cf[6 * nxy + p] * s;   // dsigma/dz = 1/d
}

// Free-surface right-hand side (rhsFreeSurface3D.f90:13-28, rhsLinearFreeSurface3D.f90:11-12) fused with
// the RK4 stage combination (Runge_Kutta_4.f90:38-48, 102-113, 159-170, 230-256), corner zeroing
// (Zero_Corners.f90) and damping zones (RelaxationModule_new.f90).
struct Zone { int axis, i0, i1, j0, j1, off; };
struct StageArgs {
    int nxg, nyg, nxy, dim3, nonlinear, stage, nzones, zero_corners;
    double g, dt;
    const double *E, *P;            // state at t_n
    const double *Es;               // stage surface elevation (for the rhs)
    const double *W, *Ex, *Ey, *Px, *Py;
    double *KE, *KP;                // accumulated k1 + 2 k2 + 2 k3
    double *Eo, *Po;                // output: next stage state, or the new state at stage 4
    const Zone *zones; const double *gam;
};
__global__ void k_stage(StageArgs a) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.nxy) return;
    const int i = p % a.nxg, j = p / a.nxg;
    double kE, kP;
    const double w = a.W[p], e = a.Es[p];
    if (!a.nonlinear) { kE = w; kP = -a.g * e; }
    else if (!a.dim3) {
        const double ex = a.Ex[p], px = a.Px[p];
        kE = -ex * px + w * (1.0 + ex * ex);
        kP = -a.g * e - 0.5 * (px * px - w * w * (1.0 + ex * ex));
    } else {
        const double ex = a.Ex[p], ey = a.Ey[p], px = a.Px[p], py = a.Py[p];
        kE = -ex * px - ey * py + w * (1.0 + ex * ex + ey * ey);
        kP = -a.g * e - 0.5 * (px * px + py * py - w * w * (1.0 + ex * ex + ey * ey));
    }
    if (a.zero_corners && a.dim3 && (i == 0 || i == a.nxg - 1) && (j == 0 || j == a.nyg - 1)) { kE = 0.0; kP = 0.0; }
    double ne, np_;
    if (a.stage == 1) { a.KE[p] = kE; a.KP[p] = kP; ne = a.E[p] + 0.5 * a.dt * kE; np_ = a.P[p] + 0.5 * a.dt * kP; }
    else if (a.stage == 2) { a.KE[p] += 2.0 * kE; a.KP[p] += 2.0 * kP; ne = a.E[p] + 0.5 * a.dt * kE; np_ = a.P[p] + 0.5 * a.dt * kP; }
    else if (a.stage == 3) { a.KE[p] += 2.0 * kE; a.KP[p] += 2.0 * kP; ne = a.E[p] + a.dt * kE; np_ = a.P[p] + a.dt * kP; }
    else { ne = a.E[p] + a.dt / 6.0 * (a.KE[p] + kE); np_ = a.P[p] + a.dt / 6.0 * (a.KP[p] + kP); }
    for (int z = 0; z < a.nzones; z++) {
        const Zone Z = a.zones[z];
        if (i >= Z.i0 && i <= Z.i1 && j >= Z.j0 && j <= Z.j1) {
            const double gm = a.gam[Z.off + (Z.axis == 0 ? i - Z.i0 : j - Z.j0)];
            ne = ne * gm; np_ = np_ * gm;
        }
    }
    a.Eo[p] = ne; a.Po[p] = np_;
}

// Laplace right-hand side: Dirichlet data P on the free surface of non-ghost columns, zero elsewhere
// (Runge_Kutta_4.f90:65-69 with no wall flux).
__global__ void k_fill_rhs(int nxg, int nyg, int nzg, int gx, int gy, const double *__restrict__ P, double *__restrict__ b) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)nxg * nyg * nzg;
    if (q >= n) return;
    const int k = (int)(q % nzg);
    const size_t col = q / nzg;
    const int i = (int)(col % nxg), j = (int)(col / nxg);
    double v = 0.0;
    if (k == nzg - 1 && i >= gx && i < nxg - gx && j >= gy && j < nyg - gy) v = P[col];
    b[q] = v;
}

// ---------------------------------------------------------------------------------------------
// Vector kernels (grid-stride)
// ---------------------------------------------------------------------------------------------
// y = a*x + b*y
__global__ void k_axpby(size_t n, double a, const double *__restrict__ x, double b, double *__restrict__ y) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        y[q] = a * x[q] + b * y[q];
}
// y = b - y
__global__ void k_bminus(size_t n, const double *__restrict__ b, double *__restrict__ y) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        y[q] = b[q] - y[q];
}
template <typename S, typename D>
__global__ void k_convert(size_t n, const S *__restrict__ x, D *__restrict__ y) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        y[q] = (D)x[q];
}
template <typename T>
__global__ void k_fill(size_t n, T v, T *__restrict__ y) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        y[q] = v;
}

// Block reduction of MAXK accumulators into part[blockIdx.x*MAXK + i].
template <int MAXK>
__device__ __forceinline__ void block_reduce_store(double (&acc)[MAXK], int kk, double *part) {
    __shared__ double red[MAXK][4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < MAXK; i++) {
        double v = acc[i];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) red[i][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < kk) {
        const int i = threadIdx.x;
        part[(size_t)blockIdx.x * MAXK + i] = ((red[i][0] + red[i][1]) + red[i][2]) + red[i][3];
    }
}

// part[b][i] = sum over block b's range of V_i . w, i < k  (V_i = V + i*n); with_norm adds w . w at index k.
template <int MAXK>
__global__ __launch_bounds__(256) void k_mdot(size_t n, int k, const double *__restrict__ V, const double *__restrict__ w, double *part,
                                               int with_norm) {
    double acc[MAXK];
#pragma unroll
    for (int i = 0; i < MAXK; i++) acc[i] = 0.0;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const double wq = w[q];
#pragma unroll
        for (int i = 0; i < MAXK; i++) {
            if (i < k) acc[i] += V[(size_t)i * n + q] * wq;
            else if (i == k && with_norm) acc[i] += wq * wq;
        }
    }
    block_reduce_store<MAXK>(acc, k + (with_norm ? 1 : 0), part);
}

// v <- v - sum_i h_i V_i; part gets the dots of the new v with V_i (i < k) and its squared norm (index k).
template <int MAXK>
__global__ __launch_bounds__(256) void k_gs_update_dot(size_t n, int k, const double *__restrict__ V, const double *__restrict__ h,
                                                        double *__restrict__ v, double *part) {
    double acc[MAXK];
    double hh[MAXK];
#pragma unroll
    for (int i = 0; i < MAXK; i++) { acc[i] = 0.0; hh[i] = (i < k) ? h[i] : 0.0; }
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        double x = v[q];
#pragma unroll
        for (int i = 0; i < MAXK; i++) if (i < k) x -= hh[i] * V[(size_t)i * n + q];
        v[q] = x;
#pragma unroll
        for (int i = 0; i < MAXK; i++) {
            if (i < k) acc[i] += V[(size_t)i * n + q] * x;
            else if (i == k) acc[i] += x * x;
        }
    }
    block_reduce_store<MAXK>(acc, k + 1, part);
}

// v <- (v - sum_i h_i V_i) * scale, plus the solution update form x <- x + sum_i h_i V_i (sign = +1).
template <int MAXK>
__global__ __launch_bounds__(256) void k_gs_update(size_t n, int k, const double *__restrict__ V, const double *__restrict__ h,
                                                    double sign, double scale, double *__restrict__ v) {
    double hh[MAXK];
#pragma unroll
    for (int i = 0; i < MAXK; i++) hh[i] = (i < k) ? h[i] : 0.0;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        double x = v[q];
#pragma unroll
        for (int i = 0; i < MAXK; i++) if (i < k) x += sign * hh[i] * V[(size_t)i * n + q];
        v[q] = x * scale;
    }
}

// out[i] = sum_b part[b][i]; one block per entry, fixed summation order (deterministic).
template <int MAXK>
__global__ __launch_bounds__(256) void k_reduce_parts(int nblk, int k, const double *part, double *out) {
    const int i = blockIdx.x;
    if (i >= k) return;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 256) s += part[(size_t)b * MAXK + i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    __shared__ double w[4];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[i] = ((w[0] + w[1]) + w[2]) + w[3];
}

// ---------------------------------------------------------------------------------------------
// Multigrid (low-order linearised operator, matrix-free; BuildLinearSystemMatrix.f90 row semantics)
// ---------------------------------------------------------------------------------------------
template <typename T>
struct LvArgs {
    int nxg, nyg, nzg, nxy, gx, gy, dim3;
    const T *sx, *sy, *sz;  // 3-point stencil rows [N][2][3]
    const T *zs;            // sigma of each level (0 on the ghost level)
    const T *col;           // 4 planes: h, h_x, h_y, h_xx+h_yy
};

template <typename T>
__device__ __forceinline__ void low_coefs(const LvArgs<T> &L, int p, int k, T &s2, T &s3, T &s4, T &s5) {
    const T d = L.col[p], hx = L.col[L.nxy + p], hy = L.col[2 * L.nxy + p], hs = L.col[3 * L.nxy + p];
    if (L.dim3) sigma_coef3<T>(L.zs[k], d, hx, hy, hs, T(0), T(0), T(0), s2, s3, s4, s5);
    else { sigma_coef2<T>(L.zs[k], d, hx, hs, T(0), T(0), s2, s3, s5); s4 = T(0); }
}

// Zebra z-line relaxation of the columns of one colour ((i+j) mod 2 == colour): exact tridiagonal
// solve in sigma with the neighbouring columns frozen.
template <typename T, int NZ>
__global__ __launch_bounds__(256) void k_mg_zebra(LvArgs<T> L, int colour, T *__restrict__ x, const T *__restrict__ b) {
    const int nxg = L.nxg, nyg = L.nyg;
    const int nxi = nxg - 2 * L.gx;
    const int j = L.dim3 ? (int)blockIdx.y + L.gy : 0;
    if (j >= nyg - L.gy) return;
    const int ifirst = L.gx + (((L.gx + j) & 1) != colour ? 1 : 0);
    const int i = ifirst + 2 * (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= L.gx + nxi) return;
    const int p = i + nxg * j;
    auto X = [&](int k, int ii, int jj) -> T { return x[((size_t)jj * nxg + ii) * NZ + k]; };
    const T *bc = b + (size_t)p * NZ;
    const int wsx = cclamp(i - 1, 0, nxg - 3), wsy = L.dim3 ? cclamp(j - 1, 0, nyg - 3) : 0;
    const T *SX1 = L.sx + (i * 2) * 3, *SX2 = SX1 + 3;
    const T *SY1 = L.sy + (j * 2) * 3, *SY2 = SY1 + 3;
    int shx = 0; if (wsx == 0) shx = 1; if (wsx + 2 == nxg - 1) shx = -1;
    int shy = 0;
    if (L.dim3) { if (wsy == 0) shy = 1; if (wsy + 2 == nyg - 1) shy = -1; }
    const T *SXc = L.sx + ((i - shx) * 2) * 3;
    const T *SYc = L.sy + ((j - shy) * 2) * 3;
    T lo[NZ], di[NZ], up[NZ], rh[NZ];
    T e2 = T(0);
#pragma unroll
    for (int k = 0; k < NZ; k++) { lo[k] = T(0); di[k] = T(0); up[k] = T(0); }
    // bottom ghost row
    {
        const T *Sz = L.sz + (1 * 2) * 3;
        T s2, s3, s4, s5;
        low_coefs(L, p, 1, s2, s3, s4, s5);
        const T hx = L.col[L.nxy + p], hy = L.dim3 ? L.col[2 * L.nxy + p] : T(0);
        const T cz = s5 + hx * s2 + hy * s4;
        T a0 = cz * Sz[0], a1 = cz * Sz[1], a2 = cz * Sz[2], off = T(0);
        for (int m = 0; m < 3; m++) {
            if (wsx + m == i) a1 += hx * SX1[m]; else off += hx * SX1[m] * X(1, wsx + m, j);
        }
        if (L.dim3)
            for (int m = 0; m < 3; m++) {
                if (wsy + m == j) a1 += hy * SY1[m]; else off += hy * SY1[m] * X(1, i, wsy + m);
            }
        di[0] = a0; up[0] = a1; e2 = a2;
        rh[0] = bc[0] - off;
    }
#pragma unroll
    for (int k = 1; k < NZ - 1; k++) {
        const T *Z1 = L.sz + (k * 2) * 3, *Z2 = Z1 + 3;
        T s2, s3, s4, s5;
        low_coefs(L, p, k, s2, s3, s4, s5);
        const T czz = s2 * s2 + s4 * s4 + s5 * s5;
        T a[3];
        for (int m = 0; m < 3; m++) a[m] = s3 * Z1[m] + czz * Z2[m];
        T off = T(0);
        for (int m = 0; m < 3; m++) {
            if (wsx + m == i) a[1] += SX2[m]; else off += SX2[m] * X(k, wsx + m, j);
        }
        for (int mz = 0; mz < 3; mz++) {
            const int kk = k - 1 + mz;
            const int sh = (kk == 0) ? shx : 0;
            const T *S1 = (kk == 0) ? SXc : SX1;
            for (int m = 0; m < 3; m++) {
                const T w = T(2) * s2 * Z1[mz] * S1[m];
                const int ii = wsx + sh + m;
                if (ii == i) a[mz] += w;
                else if (ii - i != 2 && i - ii != 2) off += w * X(kk, ii, j);  // same-colour column: not coupled in the sweep
            }
        }
        if (L.dim3) {
            for (int m = 0; m < 3; m++) {
                if (wsy + m == j) a[1] += SY2[m]; else off += SY2[m] * X(k, i, wsy + m);
            }
            for (int mz = 0; mz < 3; mz++) {
                const int kk = k - 1 + mz;
                const int sh = (kk == 0) ? shy : 0;
                const T *S1 = (kk == 0) ? SYc : SY1;
                for (int m = 0; m < 3; m++) {
                    const T w = T(2) * s4 * Z1[mz] * S1[m];
                    const int jj = wsy + sh + m;
                    if (jj == j) a[mz] += w;
                    else if (jj - j != 2 && j - jj != 2) off += w * X(kk, i, jj);  // same-colour column: not coupled in the sweep
                }
            }
        }
        lo[k] = a[0]; di[k] = a[1]; up[k] = a[2];
        rh[k] = bc[k] - off;
    }
    di[NZ - 1] = T(1); rh[NZ - 1] = bc[NZ - 1];
    if (e2 != T(0)) {
        const T f = e2 / up[1];
        di[0] -= f * lo[1]; up[0] -= f * di[1]; rh[0] -= f * rh[1];
    }
#pragma unroll
    for (int k = 1; k < NZ; k++) {
        const T f = lo[k] / di[k - 1];
        di[k] -= f * up[k - 1];
        rh[k] -= f * rh[k - 1];
    }
    T *xc = x + (size_t)p * NZ;
    T xn[NZ];
    xn[NZ - 1] = rh[NZ - 1] / di[NZ - 1];
#pragma unroll
    for (int k = NZ - 2; k >= 0; k--) xn[k] = (rh[k] - up[k] * xn[k + 1]) / di[k];
#pragma unroll
    for (int k = 0; k < NZ; k++) xc[k] = xn[k];
}

// LDS-tiled zebra z-line relaxation: a block stages a 32 x 8 tile of columns (plus a one-column
// halo) and the tile's right-hand sides, 128 threads each solve the sweep-colour column of one
// (even, odd) pair, and the whole tile is written back coalesced (columns of the other colour
// are rewritten unchanged; nothing else modifies them during the sweep).
template <typename T, int NZ>
__global__ __launch_bounds__(128) void k_mg_zebra_t(LvArgs<T> L, int colour, T *__restrict__ x, const T *__restrict__ b) {
    constexpr int TX = 32, TY = 8, WX = TX + 2, WY = TY + 2, NZP = NZ | 1;
    __shared__ T sX[WX * WY * NZP];
    __shared__ T sB[TX * TY * NZP];
    const int nxg = L.nxg, nyg = L.nyg, d3 = L.dim3;
    const int i0 = L.gx + (int)blockIdx.x * TX, j0 = d3 ? L.gy + (int)blockIdx.y * TY : 0;
    const int ie = cmin(i0 + TX, nxg - L.gx), je = d3 ? cmin(j0 + TY, nyg - L.gy) : 1;
    const int lox = i0 - 1, loy = d3 ? j0 - 1 : 0;
    const int wcols = cmin(WX, nxg - lox), wrows = d3 ? cmin(WY, nyg - loy) : 1;
    for (int r = 0; r < wrows; r++) {
        const T *src = x + ((size_t)(loy + r) * nxg + lox) * NZ;
        for (int e = threadIdx.x; e < wcols * NZ; e += blockDim.x) {
            const int c = e / NZ, k = e - c * NZ;
            sX[(r * WX + c) * NZP + k] = src[e];
        }
    }
    for (int r = 0; r < je - j0; r++) {
        const T *src = b + ((size_t)(j0 + r) * nxg + i0) * NZ;
        for (int e = threadIdx.x; e < (ie - i0) * NZ; e += blockDim.x) {
            const int c = e / NZ, k = e - c * NZ;
            sB[(r * TX + c) * NZP + k] = src[e];
        }
    }
    __syncthreads();
    auto X = [&](int k, int ii, int jj) -> T { return sX[((jj - loy) * WX + (ii - lox)) * NZP + k]; };
    const int ty = threadIdx.x / 16, pr = threadIdx.x % 16;
    const int j = j0 + ty;
    int i = i0 + 2 * pr;
    if (((i + j) & 1) != colour) i++;
    if (j < je && i < ie) {
        const int p = i + nxg * j;
        const T *bc = sB + ((j - j0) * TX + (i - i0)) * NZP;
        const int wsx = cclamp(i - 1, 0, nxg - 3), wsy = d3 ? cclamp(j - 1, 0, nyg - 3) : 0;
        const T *SX1 = L.sx + (i * 2) * 3, *SX2 = SX1 + 3;
        const T *SY1 = L.sy + (j * 2) * 3, *SY2 = SY1 + 3;
        int shx = 0; if (wsx == 0) shx = 1; if (wsx + 2 == nxg - 1) shx = -1;
        int shy = 0;
        if (d3) { if (wsy == 0) shy = 1; if (wsy + 2 == nyg - 1) shy = -1; }
        const T *SXc = L.sx + ((i - shx) * 2) * 3;
        const T *SYc = L.sy + ((j - shy) * 2) * 3;
        T lo[NZ], di[NZ], up[NZ], rh[NZ];
        T e2 = T(0);
#pragma unroll
        for (int k = 0; k < NZ; k++) { lo[k] = T(0); di[k] = T(0); up[k] = T(0); }
        {
            const T *Sz = L.sz + (1 * 2) * 3;
            T s2, s3, s4, s5;
            low_coefs(L, p, 1, s2, s3, s4, s5);
            const T hx = L.col[L.nxy + p], hy = d3 ? L.col[2 * L.nxy + p] : T(0);
            const T cz = s5 + hx * s2 + hy * s4;
            T a0 = cz * Sz[0], a1 = cz * Sz[1], a2 = cz * Sz[2], off = T(0);
            for (int m = 0; m < 3; m++) {
                if (wsx + m == i) a1 += hx * SX1[m]; else off += hx * SX1[m] * X(1, wsx + m, j);
            }
            if (d3)
                for (int m = 0; m < 3; m++) {
                    if (wsy + m == j) a1 += hy * SY1[m]; else off += hy * SY1[m] * X(1, i, wsy + m);
                }
            di[0] = a0; up[0] = a1; e2 = a2;
            rh[0] = bc[0] - off;
        }
#pragma unroll
        for (int k = 1; k < NZ - 1; k++) {
            const T *Z1 = L.sz + (k * 2) * 3, *Z2 = Z1 + 3;
            T s2, s3, s4, s5;
            low_coefs(L, p, k, s2, s3, s4, s5);
            const T czz = s2 * s2 + s4 * s4 + s5 * s5;
            T a[3];
            for (int m = 0; m < 3; m++) a[m] = s3 * Z1[m] + czz * Z2[m];
            T off = T(0);
            for (int m = 0; m < 3; m++) {
                if (wsx + m == i) a[1] += SX2[m]; else off += SX2[m] * X(k, wsx + m, j);
            }
            for (int mz = 0; mz < 3; mz++) {
                const int kk = k - 1 + mz;
                const int sh = (kk == 0) ? shx : 0;
                const T *S1 = (kk == 0) ? SXc : SX1;
                for (int m = 0; m < 3; m++) {
                    const T w = T(2) * s2 * Z1[mz] * S1[m];
                    const int ii = wsx + sh + m;
                    if (ii == i) a[mz] += w;
                    else if (ii - i != 2 && i - ii != 2) off += w * X(kk, ii, j);
                }
            }
            if (d3) {
                for (int m = 0; m < 3; m++) {
                    if (wsy + m == j) a[1] += SY2[m]; else off += SY2[m] * X(k, i, wsy + m);
                }
                for (int mz = 0; mz < 3; mz++) {
                    const int kk = k - 1 + mz;
                    const int sh = (kk == 0) ? shy : 0;
                    const T *S1 = (kk == 0) ? SYc : SY1;
                    for (int m = 0; m < 3; m++) {
                        const T w = T(2) * s4 * Z1[mz] * S1[m];
                        const int jj = wsy + sh + m;
                        if (jj == j) a[mz] += w;
                        else if (jj - j != 2 && j - jj != 2) off += w * X(kk, i, jj);
                    }
                }
            }
            lo[k] = a[0]; di[k] = a[1]; up[k] = a[2];
            rh[k] = bc[k] - off;
        }
        di[NZ - 1] = T(1); rh[NZ - 1] = bc[NZ - 1];
        if (e2 != T(0)) {
            const T f = e2 / up[1];
            di[0] -= f * lo[1]; up[0] -= f * di[1]; rh[0] -= f * rh[1];
        }
#pragma unroll
        for (int k = 1; k < NZ; k++) {
            const T f = lo[k] / di[k - 1];
            di[k] -= f * up[k - 1];
            rh[k] -= f * rh[k - 1];
        }
        T *own = sX + ((j - loy) * WX + (i - lox)) * NZP;
        T xn = rh[NZ - 1] / di[NZ - 1];
        own[NZ - 1] = xn;
#pragma unroll
        for (int k = NZ - 2; k >= 0; k--) { xn = (rh[k] - up[k] * xn) / di[k]; own[k] = xn; }
    }
    __syncthreads();
    for (int r = 0; r < je - j0; r++) {
        T *dst = x + ((size_t)(j0 + r) * nxg + i0) * NZ;
        for (int e = threadIdx.x; e < (ie - i0) * NZ; e += blockDim.x) {
            const int c = e / NZ, k = e - c * NZ;
            dst[e] = sX[((j0 + r - loy) * WX + (i0 + c - lox)) * NZP + k];
        }
    }
}

// Wall ghost points, one thread per (ghost column, level): Neumann rows relaxed pointwise, identity
// rows set to the right-hand side (the role of GaussSeidelSpecialPoints.f90).
template <typename T>
__global__ void k_mg_ghost_p(LvArgs<T> L, T *__restrict__ x, const T *__restrict__ b) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nxg = L.nxg, nyg = L.nyg, nzg = L.nzg;
    const int col = t / nzg, k = t - col * nzg;
    int i, j;
    if (L.dim3) {
        const int nb = 2 * nxg + 2 * (nyg - 2);
        if (col >= nb) return;
        if (col < nxg) { i = col; j = 0; }
        else if (col < 2 * nxg) { i = col - nxg; j = nyg - 1; }
        else if (col < 2 * nxg + nyg - 2) { i = 0; j = 1 + col - 2 * nxg; }
        else { i = nxg - 1; j = 1 + col - 2 * nxg - (nyg - 2); }
    } else {
        if (col >= 2) return;
        i = col == 0 ? 0 : nxg - 1; j = 0;
    }
    const size_t q = ((size_t)j * nxg + i) * nzg + k;
    const bool wx = L.gx && (i == 0 || i == nxg - 1), wy = L.gy && (j == 0 || j == nyg - 1);
    if ((wx && wy) || k == 0) { x[q] = b[q]; return; }
    T s = T(0), dg = T(1);
    if (wx) {
        const int ib = i == 0 ? 1 : nxg - 2, ws = cclamp(ib - 1, 0, nxg - 3);
        const T *S = L.sx + (ib * 2) * 3;
        for (int m = 0; m < 3; m++) {
            if (ws + m == i) dg = S[m];
            else s += S[m] * x[((size_t)j * nxg + ws + m) * nzg + k];
        }
    } else {
        const int jb = j == 0 ? 1 : nyg - 2, ws = cclamp(jb - 1, 0, nyg - 3);
        const T *S = L.sy + (jb * 2) * 3;
        for (int m = 0; m < 3; m++) {
            if (ws + m == j) dg = S[m];
            else s += S[m] * x[((size_t)(ws + m) * nxg + i) * nzg + k];
        }
    }
    x[q] = (b[q] - s) / dg;
}

// LDS-tiled residual r = b - L x (16 x 16 columns per block, halo 2 on domain ends as the matvec).
template <typename T, int NZ>
__global__ __launch_bounds__(256) void k_mg_residual_t(LvArgs<T> L, const T *__restrict__ x, const T *__restrict__ b, T *__restrict__ r) {
    constexpr int TX = 16, TY = 16, R = 2, WX = TX + 2 * R, WY = TY + 2 * R, NZP = NZ | 1;
    __shared__ T sX[WX * WY * NZP];
    const int nxg = L.nxg, nyg = L.nyg, d3 = L.dim3;
    const int tyd = d3 ? TY : 1;
    const int i0 = d3 ? (int)blockIdx.x * TX : (int)blockIdx.x * (TX * TY), j0 = d3 ? (int)blockIdx.y * TY : 0;
    const int tw = d3 ? TX : TX * TY;        // columns per tile row
    const int lox = cmax(cmin(i0 - R, nxg - 4), 0);
    const int loy = d3 ? cmax(cmin(j0 - R, nyg - 4), 0) : 0;
    const int wxmax = d3 ? WX : WX * WY;     // 2D: the tile is one long row
    const int wcols = cmin(tw + 2 * R, nxg - lox), wrows = d3 ? cmin(WY, nyg - loy) : 1;
    (void)wxmax;
    const int rowstride = d3 ? WX : (TX * TY + 2 * R);
    for (int rr = 0; rr < wrows; rr++) {
        const T *src = x + ((size_t)(loy + rr) * nxg + lox) * NZ;
        for (int e = threadIdx.x; e < wcols * NZ; e += blockDim.x) {
            const int c = e / NZ, k = e - c * NZ;
            sX[(rr * rowstride + c) * NZP + k] = src[e];
        }
    }
    __syncthreads();
    auto X = [&](int k, int ii, int jj) -> T { return sX[((jj - loy) * rowstride + (ii - lox)) * NZP + k]; };
    const int tx = d3 ? (int)threadIdx.x % TX : (int)threadIdx.x, ty = d3 ? (int)threadIdx.x / TX : 0;
    const int i = i0 + tx, j = j0 + ty;
    const bool active = i < nxg && j < nyg && (d3 || ty == 0);
    T o[NZ];
    if (active) {
        const int p = i + nxg * j;
        const T *bc = b + (size_t)p * NZ;
        const bool wx = L.gx && (i == 0 || i == nxg - 1), wy = L.gy && (j == 0 || j == nyg - 1);
        if (wx && wy) {
#pragma unroll
            for (int k = 0; k < NZ; k++) o[k] = bc[k] - X(k, i, j);
        } else if (wx || wy) {
            o[0] = bc[0] - X(0, i, j);
            if (wx) {
                const int ib = i == 0 ? 1 : nxg - 2, ws = cclamp(ib - 1, 0, nxg - 3);
                const T *S = L.sx + (ib * 2) * 3;
#pragma unroll
                for (int k = 1; k < NZ; k++) o[k] = bc[k] - ((S[0] * X(k, ws, j) + S[1] * X(k, ws + 1, j)) + S[2] * X(k, ws + 2, j));
            } else {
                const int jb = j == 0 ? 1 : nyg - 2, ws = cclamp(jb - 1, 0, nyg - 3);
                const T *S = L.sy + (jb * 2) * 3;
#pragma unroll
                for (int k = 1; k < NZ; k++) o[k] = bc[k] - ((S[0] * X(k, i, ws) + S[1] * X(k, i, ws + 1)) + S[2] * X(k, i, ws + 2));
            }
        } else {
            const int wsx = cclamp(i - 1, 0, nxg - 3), wsy = d3 ? cclamp(j - 1, 0, nyg - 3) : 0;
            const T *SX1 = L.sx + (i * 2) * 3, *SX2 = SX1 + 3;
            const T *SY1 = L.sy + (j * 2) * 3, *SY2 = SY1 + 3;
            T xc[NZ], dxl[NZ], dyl[NZ];
#pragma unroll
            for (int k = 0; k < NZ; k++) {
                xc[k] = X(k, i, j);
                dxl[k] = (SX1[0] * X(k, wsx, j) + SX1[1] * X(k, wsx + 1, j)) + SX1[2] * X(k, wsx + 2, j);
                dyl[k] = d3 ? (SY1[0] * X(k, i, wsy) + SY1[1] * X(k, i, wsy + 1)) + SY1[2] * X(k, i, wsy + 2) : T(0);
            }
            {
                int sh = 0; if (wsx == 0) sh = 1; if (wsx + 2 == nxg - 1) sh = -1;
                if (sh) { const T *S = L.sx + ((i - sh) * 2) * 3; dxl[0] = (S[0] * X(0, wsx + sh, j) + S[1] * X(0, wsx + sh + 1, j)) + S[2] * X(0, wsx + sh + 2, j); }
                if (d3) {
                    sh = 0; if (wsy == 0) sh = 1; if (wsy + 2 == nyg - 1) sh = -1;
                    if (sh) { const T *S = L.sy + ((j - sh) * 2) * 3; dyl[0] = (S[0] * X(0, i, wsy + sh) + S[1] * X(0, i, wsy + sh + 1)) + S[2] * X(0, i, wsy + sh + 2); }
                }
            }
            o[NZ - 1] = bc[NZ - 1] - xc[NZ - 1];
#pragma unroll
            for (int k = 1; k < NZ - 1; k++) {
                const T *Z1 = L.sz + (k * 2) * 3, *Z2 = Z1 + 3;
                const T dz = (Z1[0] * xc[k - 1] + Z1[1] * xc[k]) + Z1[2] * xc[k + 1];
                const T dzz = (Z2[0] * xc[k - 1] + Z2[1] * xc[k]) + Z2[2] * xc[k + 1];
                const T dxx = (SX2[0] * X(k, wsx, j) + SX2[1] * X(k, wsx + 1, j)) + SX2[2] * X(k, wsx + 2, j);
                const T dyy = d3 ? (SY2[0] * X(k, i, wsy) + SY2[1] * X(k, i, wsy + 1)) + SY2[2] * X(k, i, wsy + 2) : T(0);
                const T dxz = (Z1[0] * dxl[k - 1] + Z1[1] * dxl[k]) + Z1[2] * dxl[k + 1];
                const T dyz = (Z1[0] * dyl[k - 1] + Z1[1] * dyl[k]) + Z1[2] * dyl[k + 1];
                T s2, s3, s4, s5;
                low_coefs(L, p, k, s2, s3, s4, s5);
                o[k] = bc[k] - (dxx + dyy + s3 * dz + T(2) * (s2 * dxz + s4 * dyz) + (s2 * s2 + s4 * s4 + s5 * s5) * dzz);
            }
            {
                const T *Z1 = L.sz + (1 * 2) * 3;
                const T dz = (Z1[0] * xc[0] + Z1[1] * xc[1]) + Z1[2] * xc[2];
                T s2, s3, s4, s5;
                low_coefs(L, p, 1, s2, s3, s4, s5);
                const T hx = L.col[L.nxy + p], hy = d3 ? L.col[2 * L.nxy + p] : T(0);
                o[0] = bc[0] - ((s5 + hx * s2 + hy * s4) * dz + hx * dxl[1] + hy * dyl[1]);
            }
        }
    }
    __syncthreads();
    if (active) {
        T *dst = sX + ((j - loy) * rowstride + (i - lox)) * NZP;
#pragma unroll
        for (int k = 0; k < NZ; k++) dst[k] = o[k];
    }
    __syncthreads();
    const int ie = cmin(i0 + tw, nxg), je = d3 ? cmin(j0 + TY, nyg) : 1;
    for (int rr = 0; rr < je - j0; rr++) {
        T *dst = r + ((size_t)(j0 + rr) * nxg + i0) * NZ;
        for (int e = threadIdx.x; e < (ie - i0) * NZ; e += blockDim.x) {
            const int c = e / NZ, k = e - c * NZ;
            dst[e] = sX[((j0 + rr - loy) * rowstride + (i0 + c - lox)) * NZP + k];
        }
    }
}

// Restriction (MGRestriction3DExplicit.f90): separable full weighting x -> y -> z on interior
// points below the free surface, then injection; coarse ghost entries are zero.
struct RestrictGeo {
    int nxgf, nygf, nzgf, nxf, nyf, nzf, gx, gy;     // fine
    int nxgc, nygc, nzgc, nxc, nyc, nzc;             // coarse
    int xc, yc, zc;
    const double *xf, *yf, *zf;                      // fine coordinates
};
template <typename T>
__global__ void k_mg_restrict(RestrictGeo G, const T *__restrict__ rf, T *__restrict__ rc) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nc = (size_t)G.nxgc * G.nygc * G.nzgc;
    if (q >= nc) return;
    const int kc = (int)(q % G.nzgc);
    const size_t col = q / G.nzgc;
    const int ic = (int)(col % G.nxgc), jc = (int)(col / G.nxgc);
    const int icn = ic - G.gx, jcn = jc - G.gy;
    if (kc < 1 || icn < 0 || icn >= G.nxc || jcn < 0 || jcn >= G.nyc) { rc[q] = T(0); return; }
    const int kf = 1 + (kc - 1) * (1 + G.zc), ifn = G.gx + icn * (1 + G.xc), jfn = G.gy + jcn * (1 + G.yc);
    auto R = [&](int k, int i, int j) -> T { return rf[((size_t)j * G.nxgf + i) * G.nzgf + k]; };
    // x pass value at (k,i,j)
    auto Xv = [&](int k, int i, int j) -> T {
        if (G.xc && k >= 1 && k <= G.nzf - 1 && j >= G.gy && j <= G.nyf + G.gy - 1 && ((i - G.gx) % 2 == 0) && i >= G.gx && i <= G.nxf + G.gx - 1) {
            const T dW = (T)(G.xf[i] - G.xf[i - 1]), dE = (T)(G.xf[i + 1] - G.xf[i]);
            const T dxc = T(1) / (dW + dE);
            return (T(0.5) * dxc) * ((dW * R(k, i - 1, j) + (dW + dE) * R(k, i, j)) + dE * R(k, i + 1, j));
        }
        return R(k, i, j);
    };
    auto Yv = [&](int k, int i, int j) -> T {
        if (G.yc && k >= 1 && k <= G.nzf - 1 && ((j - G.gy) % 2 == 0) && j >= G.gy && j <= G.nyf + G.gy - 1 &&
            i >= G.gx && i <= G.nxf + G.gx - 1 && ((i - G.gx) % (1 + G.xc) == 0)) {
            const T dS = (T)(G.yf[j] - G.yf[j - 1]), dN = (T)(G.yf[j + 1] - G.yf[j]);
            const T dyc = T(1) / (dS + dN);
            return (T(0.5) * dyc) * ((dS * Xv(k, i, j - 1) + (dN + dS) * Xv(k, i, j)) + dN * Xv(k, i, j + 1));
        }
        return Xv(k, i, j);
    };
    T v;
    if (G.zc && kf >= 1 && kf <= G.nzf - 1 && ((kf - 1) % 2 == 0)) {
        const T dU = (T)(G.zf[kf + 1] - G.zf[kf]), dD = (T)(G.zf[kf] - G.zf[kf - 1]);
        const T dzc = T(1) / (dD + dU);
        v = (T(0.5) * dzc) * (dD * Yv(kf - 1, ifn, jfn) + dU * Yv(kf + 1, ifn, jfn)) + T(0.5) * Yv(kf, ifn, jfn);
    } else v = Yv(kf, ifn, jfn);
    rc[q] = v;
}

// Prolongation (MGProlongation3DExplicit.f90) of the coarse correction, added to x on interior
// columns below the free surface.
template <typename T>
__global__ void k_mg_prolong_add(RestrictGeo G, const T *__restrict__ ec, T *__restrict__ xf) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nf = (size_t)G.nxgf * G.nygf * G.nzgf;
    if (q >= nf) return;
    const int k = (int)(q % G.nzgf);
    const size_t col = q / G.nzgf;
    const int i = (int)(col % G.nxgf), j = (int)(col / G.nxgf);
    if (k < 1 || k > G.nzf - 1 || i < G.gx || i > G.nxf + G.gx - 1 || j < G.gy || j > G.nyf + G.gy - 1) return;
    const int XF = G.xc ? 2 : 1, YF = G.yc ? 2 : 1, ZF = G.zc ? 2 : 1;
    auto E0 = [&](int kk, int ii, int jj) -> T {
        if (kk < 1 || kk > G.nzf - 1 || ((kk - 1) % ZF) != 0) return T(0);
        if (((ii - G.gx) % XF) != 0 || ((jj - G.gy) % YF) != 0) return T(0);
        const int kc = 1 + (kk - 1) / ZF, ic = G.gx + (ii - G.gx) / XF, jc = G.gy + (jj - G.gy) / YF;
        if (kc > G.nzc - 1) return T(0);
        return ec[((size_t)jc * G.nxgc + ic) * G.nzgc + kc];
    };
    auto EX = [&](int kk, int ii, int jj) -> T {
        if (XF == 2 && ((jj - G.gy) % YF) == 0 && ((ii - G.gx) % 2) == 1 && kk >= 1 && kk <= G.nzf - 1 && ((kk - 1) % ZF) == 0) {
            const T dW = (T)(G.xf[ii] - G.xf[ii - 1]), dE = (T)(G.xf[ii + 1] - G.xf[ii]);
            const T dxf = T(1) / (dW + dE);
            return (dE * E0(kk, ii - 1, jj) + dW * E0(kk, ii + 1, jj)) * dxf;
        }
        return E0(kk, ii, jj);
    };
    auto EY = [&](int kk, int ii, int jj) -> T {
        if (kk > G.nzf - 1) return T(0);
        if (YF == 2 && ((jj - G.gy) % 2) == 1 && kk >= 1 && ((kk - 1) % ZF) == 0) {
            const T dS = (T)(G.yf[jj] - G.yf[jj - 1]), dN = (T)(G.yf[jj + 1] - G.yf[jj]);
            const T dyf = T(1) / (dS + dN);
            return (dN * EX(kk, ii, jj - 1) + dS * EX(kk, ii, jj + 1)) * dyf;
        }
        return EX(kk, ii, jj);
    };
    T e;
    if (ZF == 2 && ((k - 1) % 2) == 1) {
        const T dD = (T)(G.zf[k] - G.zf[k - 1]), dU = (T)(G.zf[k + 1] - G.zf[k]);
        const T dzf = T(1) / (dD + dU);
        e = (dU * EY(k - 1, i, j) + dD * EY(k + 1, i, j)) * dzf;
    } else e = EY(k, i, j);
    xf[q] = xf[q] + e;
}

// Coarsest level: x = Minv b (one block).
template <typename T>
__global__ __launch_bounds__(256) void k_mg_coarse(int n, const T *__restrict__ Minv, const T *__restrict__ b, T *__restrict__ x) {
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
        T s = T(0);
        for (int c = 0; c < n; c++) s += Minv[(size_t)r * n + c] * b[c];
        x[r] = s;
    }
}

}  // namespace ow3d
