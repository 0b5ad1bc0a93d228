// HIP kernels for gfx950 (MI355X) of the Laplace-solve-per-RK4-stage path.
//
// Layout (as the reference, BuildLinearSystem.f90:117): sigma-fastest columns,
//   q = k + nzg*(i + nxg*j);  plane fields p = i + nxg*j.
// The high-order matvec keeps whole sigma columns of a 2D tile (plus an alpha+1 halo) in LDS;
// every interior column's x-, y-, sigma- and mixed derivatives are formed from that tile, so
// phi is read from HBM once per matvec (halo re-reads hit L2 / MALL) and y written once.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>
#include <cstdint>

namespace ow3d {

__host__ __device__ constexpr int cmin(int a, int b) { return a < b ? a : b; }
__host__ __device__ constexpr int cmax(int a, int b) { return a > b ? a : b; }
__host__ __device__ constexpr int cclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Kernels launched inside a device-driven Laplace solve take the solve's "done" flag and return at once
// when it is set, so a fixed launch sequence (a captured graph) costs almost nothing after convergence.
// nullptr: always run.
__device__ __forceinline__ bool solve_done(const int *skip) { return skip != nullptr && *skip != 0; }

// ---------------------------------------------------------------------------------------------
// High-order matvec  (BuildLinearSystem.f90:1-290)
// ---------------------------------------------------------------------------------------------
struct MvArgs {
    int nxg, nyg, nxy;
    const double *sx, *sy, *sz;  // stencil rows [N][2][2A+1]
    const double *zs;            // sigma of each level (0 on the ghost level)
    const double *cf;            // 7 planes: sigma coefficients affine in sigma, s2 = A2 + sigma*B2, s4 = A4 + sigma*B4,
                                 // s3 = A3 + sigma*B3, s5 = S5; planes (A2, B2, A4, B4, A3, B3, S5)
    const int *skip;             // solve_done flag (nullptr: always run)
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

// Copies a window of rows x cols sigma-columns (NZ values each) of a field in HBM into LDS: window row r
// starts at src + r * pitch; LDS holds element (r, c, k) at (r * RS + c) * NZP + k.  All loads of a thread
// are issued before its first LDS store, so a block waits for HBM once, not once per row.
// The two halves are also available separately (WindowRegs: load into registers, store to LDS) so that a
// kernel can issue the loads of several windows before it waits for any of them.
// FP32 windows of columns with an even NZ move two values per load (8 bytes): callers pass src at an even
// element offset with an even pitch, so every pair is 8-byte aligned and stays inside one window row.
template <typename T, int NZ, int MR, int MC, int NT>
struct WindowRegs {
    static constexpr int RL = MC * NZ, TOT = MR * RL;
    static constexpr bool PAIRS = NZ % 2 == 0 && sizeof(T) == 4;
    static constexpr int W = PAIRS ? 2 : 1, PER = (TOT / W + NT - 1) / NT;
    T v[PER * W];
    __device__ __forceinline__ void load(const T *__restrict__ src, size_t pitch, int rows, int cols) {
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int e = ((int)threadIdx.x + q * NT) * W;
            const int r = e / RL, c = e - r * RL;
            const bool ok = e < TOT && r < rows && c < cols * NZ;
            if constexpr (PAIRS) {
                float2 w = make_float2(0.0f, 0.0f);
                if (ok) w = *reinterpret_cast<const float2 *>(src + (size_t)r * pitch + c);
                v[2 * q] = w.x;
                v[2 * q + 1] = w.y;
            } else {
                v[q] = ok ? src[(size_t)r * pitch + c] : T(0);
            }
        }
    }
    template <int NZP, int RS, typename U>
    __device__ __forceinline__ void store(U *lds, int rows, int cols) const {
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int e = ((int)threadIdx.x + q * NT) * W;
            const int r = e / RL, c = e - r * RL;
            if (e < TOT && r < rows && c < cols * NZ) {
                const int cc = c / NZ, k = c - cc * NZ;
                const int a = (r * RS + cc) * NZP + k;
                if constexpr (PAIRS) {
                    lds[a] = (U)v[2 * q];
                    lds[k + 1 < NZ ? a + 1 : a + NZP - NZ + 1] = (U)v[2 * q + 1];   // next level or next column
                } else {
                    lds[a] = (U)v[q];
                }
            }
        }
    }
};
// Copies one tile row of ncol columns from LDS (column c's NZ values at src + c * NZP) to dst (NZ values
// per column, contiguous), converting to O.  For an even NZ two values go per store: dst is then at an even
// element offset (a row of columns of an even depth), so every pair is aligned.
template <int NZ, int NZP, int NT, typename T, typename O>
__device__ __forceinline__ void store_row(O *__restrict__ dst, const T *src, int ncol) {
    if constexpr (NZ % 2 == 0 && (sizeof(O) == 4 || sizeof(O) == 8)) {
        using O2 = typename std::conditional<sizeof(O) == 4, float2, double2>::type;
        for (int e = 2 * (int)threadIdx.x; e < ncol * NZ; e += 2 * NT) {
            const int c = e / NZ, k = e - c * NZ, a = c * NZP + k;
            O2 w;
            w.x = (O)src[a];
            w.y = (O)src[k + 1 < NZ ? a + 1 : a + NZP - NZ + 1];   // next level or next column
            *reinterpret_cast<O2 *>(dst + e) = w;
        }
    } else {
        for (int e = (int)threadIdx.x; e < ncol * NZ; e += NT) {
            const int c = e / NZ, k = e - c * NZ;
            dst[e] = (O)src[c * NZP + k];
        }
    }
}
template <typename T, int NZ, int NZP, int MR, int MC, int RS, int NT>
__device__ __forceinline__ void stage_window(T *lds, const T *__restrict__ src, size_t pitch, int rows, int cols) {
    WindowRegs<T, NZ, MR, MC, NT> w;
    w.load(src, pitch, rows, cols);
    w.template store<NZP, RS>(lds, rows, cols);
}

// Block index -> tile index such that each of the 8 XCDs (which receive blocks round-robin) works on a
// contiguous range of tiles, so halo columns shared by neighbouring tiles are re-read from the same L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    const int q = nb / 8, rem = nb % 8, xcd = b % 8, slot = b / 8;
    return (xcd < rem ? xcd * (q + 1) : rem * (q + 1) + (xcd - rem) * q) + slot;
}

// 3D, one thread per grid point: block = 16 x TY columns x NZ levels, NT threads.  The tile is
// stored k-fastest in LDS exactly as in HBM (row segments are straight copies), and threads are mapped
// k-fastest too, so every LDS access of a wave and every global load / store is contiguous.
// Phase 1 forms the x- and y-derivatives of every tile point (the first passes of the mixed
// derivatives go to LDS); phase 2 the sigma derivatives, the mixed derivatives and the row.
template <int NZ, int A, int TY, int NT>
__global__ __launch_bounds__(NT) void k_matvec3p(MvArgs g, const double *__restrict__ phi, double *__restrict__ out) {
    if (solve_done(g.skip)) return;
    constexpr int TX = 16, RA = 2 * A + 1, R = A + 1, WX = TX + 2 * R, WY = TY + 2 * R, NC = TX * TY, NP = NC * NZ;
    constexpr int PER = (NP + NT - 1) / NT;
    __shared__ double sP[WX * WY * NZ];
    __shared__ double sDx[NP];
    __shared__ double sDy[NP];
    __shared__ double sSX[(TX + 2) * 2 * RA];
    __shared__ double sSY[(TY + 2) * 2 * RA];
    __shared__ double sSZ[NZ * 2 * RA];
    __shared__ double sZS[NZ];
    __shared__ double sCF[7 * NC];
    const int nxg = g.nxg, nyg = g.nyg, nxy = g.nxy;
    const int gxt = (nxg + TX - 1) / TX;
    const int tile = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int i0 = (tile % gxt) * TX, j0 = (tile / gxt) * TY;
    const int lox = cmax(cmin(i0 - R, nxg - 2 - 2 * A), 0);
    const int loy = cmax(cmin(j0 - R, nyg - 2 - 2 * A), 0);
    const int wx = cmin(WX, nxg - lox), wy = cmin(WY, nyg - loy);
    {
        // coefficient planes of the tile's columns, loaded together with the phi window
        constexpr int CPT = (7 * NC + NT - 1) / NT;
        double cv[CPT];
#pragma unroll
        for (int q = 0; q < CPT; q++) {
            const int e = (int)threadIdx.x + q * NT;
            const int pl = e / NC, c = e - pl * NC, ii = i0 + c % TX, jj = j0 + c / TX;
            cv[q] = (e < 7 * NC && ii < nxg && jj < nyg) ? g.cf[(size_t)pl * nxy + ii + (size_t)nxg * jj] : 0.0;
        }
        stage_window<double, NZ, NZ, WY, WX, WX, NT>(sP, phi + ((size_t)loy * nxg + lox) * NZ, (size_t)nxg * NZ, wy, wx);
#pragma unroll
        for (int q = 0; q < CPT; q++) {
            const int e = (int)threadIdx.x + q * NT;
            if (e < 7 * NC) sCF[e] = cv[q];
        }
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
            const double A2 = sCF[col], B2 = sCF[NC + col], A4 = sCF[2 * NC + col], B4 = sCF[3 * NC + col];
            const double S5 = sCF[6 * NC + col];
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
                const double A3 = sCF[4 * NC + col], B3 = sCF[5 * NC + col];
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

// 3D, one thread per sigma column: block = 16 x TY columns.  The tile and its (alpha+1)-wide halo are
// staged level-major in LDS (sP[k*PS + column], PS odd), so the 16 lanes of a tile row read consecutive
// words at every stencil tap (conflict-free) and the column's x- and y-derivatives stay in registers for
// the mixed derivatives.  Sigma-stencil rows and levels are uniform across the block (scalar loads);
// the coefficient planes of the tile are staged with the window.  Results are written back through LDS
// as coalesced row segments.
template <int NZ, int A, int TY>
__global__ __launch_bounds__(16 * TY) void k_matvec3c(MvArgs g, const double *__restrict__ phi, double *__restrict__ out) {
    if (solve_done(g.skip)) return;
    constexpr int TX = 16, NT = TX * TY, RA = 2 * A + 1, R = A + 1, WX = TX + 2 * R, WY = TY + 2 * R;
    constexpr int PS = (WX * WY) | 1;
    __shared__ double sP[NZ * PS];
    __shared__ double sSX[(TX + 2) * 2 * RA];
    __shared__ double sSY[(TY + 2) * 2 * RA];
    const int nxg = g.nxg, nyg = g.nyg, nxy = g.nxy;
    const int gxt = (nxg + TX - 1) / TX;
    const int tile = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int i0 = (tile % gxt) * TX, j0 = (tile / gxt) * TY;
    const int lox = cmax(cmin(i0 - R, nxg - 2 - 2 * A), 0);
    const int loy = cmax(cmin(j0 - R, nyg - 2 - 2 * A), 0);
    const int wx = cmin(WX, nxg - lox), wy = cmin(WY, nyg - loy);
    {
        // every HBM load of the block first: the phi window (stored level-major) and the stencil rows of
        // its columns and rows (each thread reads its own column's coefficient planes where it uses them)
        constexpr int RL = WX * NZ, TOT = WY * RL, PER = (TOT + NT - 1) / NT;
        double v[PER];
        const double *src = phi + ((size_t)loy * nxg + lox) * NZ;
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int e = (int)threadIdx.x + q * NT, r = e / RL, c = e - r * RL;
            v[q] = (e < TOT && r < wy && c < wx * NZ) ? src[(size_t)r * nxg * NZ + c] : 0.0;
        }
        for (int e = threadIdx.x; e < (TX + 2) * 2 * RA; e += NT) {
            const int r = e / (2 * RA), ii = i0 - 1 + r;
            if (ii >= 0 && ii < nxg) sSX[e] = g.sx[(size_t)ii * 2 * RA + (e - r * 2 * RA)];
        }
        for (int e = threadIdx.x; e < (TY + 2) * 2 * RA; e += NT) {
            const int r = e / (2 * RA), jj = j0 - 1 + r;
            if (jj >= 0 && jj < nyg) sSY[e] = g.sy[(size_t)jj * 2 * RA + (e - r * 2 * RA)];
        }
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int e = (int)threadIdx.x + q * NT, r = e / RL, c = e - r * RL;
            if (e < TOT && r < wy && c < wx * NZ) {
                const int cc = c / NZ, k = c - cc * NZ;
                sP[k * PS + r * WX + cc] = v[q];
            }
        }
    }
    __syncthreads();
    auto P = [&](int k, int ii, int jj) -> double { return sP[k * PS + (jj - loy) * WX + (ii - lox)]; };
    auto SX = [&](int ii) -> const double * { return sSX + (ii - (i0 - 1)) * 2 * RA; };
    auto SY = [&](int jj) -> const double * { return sSY + (jj - (j0 - 1)) * 2 * RA; };
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int i = i0 + tx, j = j0 + ty;
    const bool active = (i < nxg) && (j < nyg);
    double o[NZ];
    if (active) {
        const bool wxl = (i == 0 || i == nxg - 1), wyl = (j == 0 || j == nyg - 1);
        if (wxl && wyl) {
#pragma unroll
            for (int k = 0; k < NZ; k++) o[k] = P(k, i, j);
        } else if (wxl || wyl) {
            // wall ghost column: first derivative across the wall at the adjacent node, identity at k = 0
            const int ib = wxl ? ((i == 0) ? 1 : nxg - 2) : ((j == 0) ? 1 : nyg - 2);
            const int ws = wxl ? cclamp(ib - A, 0, nxg - RA) : cclamp(ib - A, 0, nyg - RA);
            const double *S = wxl ? SX(ib) : SY(ib);
            o[0] = P(0, i, j);
#pragma unroll
            for (int k = 1; k < NZ; k++) {
                double s = 0.0;
#pragma unroll
                for (int m = 0; m < RA; m++) s += S[m] * (wxl ? P(k, ws + m, j) : P(k, i, ws + m));
                o[k] = s;
            }
        } else {
            double dx[NZ], dy[NZ], pc[NZ];
            const int wsx = cclamp(i - A, 0, nxg - RA), wsy = cclamp(j - A, 0, nyg - RA);
            {
                const double *S1 = SX(i), *S2 = S1 + RA;
#pragma unroll
                for (int k = 0; k < NZ; k++) {
                    double d1 = 0.0, d2 = 0.0;
#pragma unroll
                    for (int m = 0; m < RA; m++) { const double v = P(k, wsx + m, j); d1 += S1[m] * v; d2 += S2[m] * v; }
                    dx[k] = d1; o[k] = d2;
                }
            }
            {
                const double *S1 = SY(j), *S2 = S1 + RA;
#pragma unroll
                for (int k = 0; k < NZ; k++) {
                    double d1 = 0.0, d2 = 0.0;
#pragma unroll
                    for (int m = 0; m < RA; m++) { const double v = P(k, i, wsy + m); d1 += S1[m] * v; d2 += S2[m] * v; }
                    dy[k] = d1; o[k] += d2;
                }
            }
#pragma unroll
            for (int k = 0; k < NZ; k++) pc[k] = P(k, i, j);
            // corner-corrected first pass on the sigma ghost level (ConstructTableCrossDerivatives.f90:173-186)
            double dxc0 = dx[0], dyc0 = dy[0];
            {
                int sh = 0;
                if (wsx == 0) sh = 1;
                if (wsx + RA - 1 == nxg - 1) sh = -1;
                if (sh != 0) {
                    const double *S = SX(i - sh);
                    double t = 0.0;
#pragma unroll
                    for (int m = 0; m < RA; m++) t += S[m] * P(0, wsx + sh + m, j);
                    dxc0 = t;
                }
                sh = 0;
                if (wsy == 0) sh = 1;
                if (wsy + RA - 1 == nyg - 1) sh = -1;
                if (sh != 0) {
                    const double *S = SY(j - sh);
                    double t = 0.0;
#pragma unroll
                    for (int m = 0; m < RA; m++) t += S[m] * P(0, i, wsy + sh + m);
                    dyc0 = t;
                }
            }
            const double *cfp = g.cf + i + (size_t)nxg * j;
            const double A2 = cfp[0], B2 = cfp[nxy], A4 = cfp[2 * (size_t)nxy], B4 = cfp[3 * (size_t)nxy];
            const double A3 = cfp[4 * (size_t)nxy], B3 = cfp[5 * (size_t)nxy], S5 = cfp[6 * (size_t)nxy];
            double obot;
            {
                constexpr int wz = (1 - A) < 0 ? 0 : (1 - A);
                const double *Z1 = g.sz + (1 * 2) * RA;
                double dz = 0.0;
#pragma unroll
                for (int m = 0; m < RA; m++) dz += Z1[m] * pc[wz + m];
                // h_x = A2 * d and h_y = A4 * d with d = 1 / S5
                const double sg = g.zs[1], d = 1.0 / S5, hx = A2 * d, hy = A4 * d;
                obot = (S5 + hx * (A2 + sg * B2) + hy * (A4 + sg * B4)) * dz + hx * dx[1] + hy * dy[1];
            }
#pragma unroll
            for (int k = 1; k < NZ - 1; k++) {
                const int wz = cclamp(k - A, 0, NZ - RA);
                const double *Z1 = g.sz + (k * 2) * RA, *Z2 = Z1 + RA;
                double dz = 0.0, dzz = 0.0, dxz = 0.0, dyz = 0.0;
#pragma unroll
                for (int m = 0; m < RA; m++) {
                    const double z1 = Z1[m], cv = pc[wz + m];
                    dz += z1 * cv;
                    dzz += Z2[m] * cv;
                    dxz += z1 * ((wz + m == 0) ? dxc0 : dx[wz + m]);
                    dyz += z1 * ((wz + m == 0) ? dyc0 : dy[wz + m]);
                }
                const double sg = g.zs[k];
                const double s2 = A2 + sg * B2, s4 = A4 + sg * B4, s3 = A3 + sg * B3;
                o[k] = o[k] + s3 * dz + 2.0 * (s2 * dxz + s4 * dyz) + (s2 * s2 + s4 * s4 + S5 * S5) * dzz;
            }
            o[0] = obot;
            o[NZ - 1] = pc[NZ - 1];
        }
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int k = 0; k < NZ; k++) sP[k * PS + (j - loy) * WX + (i - lox)] = o[k];
    }
    __syncthreads();
    {
        const int ie = cmin(i0 + TX, nxg), je = cmin(j0 + TY, nyg);
        for (int jj = j0; jj < je; jj++) {
            double *dst = out + ((size_t)jj * nxg + i0) * NZ;
            for (int e = threadIdx.x; e < (ie - i0) * NZ; e += NT) {
                const int ii = e / NZ, k = e - ii * NZ;
                dst[e] = sP[k * PS + (jj - loy) * WX + (i0 + ii - lox)];
            }
        }
    }
}

// The column matvec with the work of each column split between two waves: a block of TX x TY columns
// has 2 TX TY threads; the first TX TY take the x-side of column t (d/dx, d2/dx2, the first sigma
// derivative of phi and the mixed x-sigma term, the bottom row), the others the y-side of column
// t - TX TY (d/dy, d2/dy2, the second sigma derivative and the mixed y-sigma term): four
// 2A+1-point stencils per level on each side.  The y-side leaves its part of every row in LDS; the x-side
// adds it.  Each thread holds half the live values of the one-thread-per-column kernel and a block has
// twice the waves.  Same terms as k_matvec3c; the x and y parts are summed in another order (results agree
// to rounding).  The live values of a column take about 210 VGPRs (two waves per SIMD), so a block of
// 4 waves lets two blocks share a CU and one block's window loads overlap the other's arithmetic.  On a
// 515 x 515 x 10 grid (6th order): 16 x 8 tiles 29.6 us, 32 x 4 32.5 us, 16 x 4 33.1 us, 16 x 16 (one
// block per CU) 36.3 us.
template <int NZ, int A, int TX = 16, int TY = 8>
__global__ __launch_bounds__(2 * TX * TY) void k_matvec3s(MvArgs g, const double *__restrict__ phi, double *__restrict__ out) {
    if (solve_done(g.skip)) return;
    constexpr int NC = TX * TY, NT = 2 * NC, RA = 2 * A + 1, R = A + 1, WX = TX + 2 * R, WY = TY + 2 * R;
    constexpr int PS = (WX * WY) | 1;
    __shared__ double sP[NZ * PS];
    __shared__ double sSX[(TX + 2) * 2 * RA];
    __shared__ double sSY[(TY + 2) * 2 * RA];
    __shared__ double sYp[NZ * NC];   // y-side part of each row: [k][column]
    const int nxg = g.nxg, nyg = g.nyg, nxy = g.nxy;
    const int gxt = (nxg + TX - 1) / TX;
    const int tile = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int i0 = (tile % gxt) * TX, j0 = (tile / gxt) * TY;
    const int lox = cmax(cmin(i0 - R, nxg - 2 - 2 * A), 0);
    const int loy = cmax(cmin(j0 - R, nyg - 2 - 2 * A), 0);
    const int wx = cmin(WX, nxg - lox), wy = cmin(WY, nyg - loy);
    {
        // window values two at a time (16-byte loads) for an even NZ: rows of columns start at even
        // offsets, so every pair is aligned and stays inside one window row
        constexpr bool PAIRS = NZ % 2 == 0;
        constexpr int W = PAIRS ? 2 : 1;
        constexpr int RL = WX * NZ, TOT = WY * RL, PER = (TOT / W + NT - 1) / NT;
        double v[PER * W];
        const double *src = phi + ((size_t)loy * nxg + lox) * NZ;
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int e = ((int)threadIdx.x + q * NT) * W, r = e / RL, c = e - r * RL;
            const bool ok = e < TOT && r < wy && c < wx * NZ;
            if constexpr (PAIRS) {
                double2 w = make_double2(0.0, 0.0);
                if (ok) w = *reinterpret_cast<const double2 *>(src + (size_t)r * nxg * NZ + c);
                v[2 * q] = w.x;
                v[2 * q + 1] = w.y;
            } else {
                v[q] = ok ? src[(size_t)r * nxg * NZ + c] : 0.0;
            }
        }
        for (int e = threadIdx.x; e < (TX + 2) * 2 * RA; e += NT) {
            const int r = e / (2 * RA), ii = i0 - 1 + r;
            if (ii >= 0 && ii < nxg) sSX[e] = g.sx[(size_t)ii * 2 * RA + (e - r * 2 * RA)];
        }
        for (int e = threadIdx.x; e < (TY + 2) * 2 * RA; e += NT) {
            const int r = e / (2 * RA), jj = j0 - 1 + r;
            if (jj >= 0 && jj < nyg) sSY[e] = g.sy[(size_t)jj * 2 * RA + (e - r * 2 * RA)];
        }
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int e = ((int)threadIdx.x + q * NT) * W, r = e / RL, c = e - r * RL;
            if (e < TOT && r < wy && c < wx * NZ) {
                const int cc = c / NZ, k = c - cc * NZ;
                if constexpr (PAIRS) {
                    sP[k * PS + r * WX + cc] = v[2 * q];
                    if (k + 1 < NZ) sP[(k + 1) * PS + r * WX + cc] = v[2 * q + 1];
                    else sP[r * WX + cc + 1] = v[2 * q + 1];   // level 0 of the next column
                } else {
                    sP[k * PS + r * WX + cc] = v[q];
                }
            }
        }
    }
    __syncthreads();
    auto P = [&](int k, int ii, int jj) -> double { return sP[k * PS + (jj - loy) * WX + (ii - lox)]; };
    auto SX = [&](int ii) -> const double * { return sSX + (ii - (i0 - 1)) * 2 * RA; };
    auto SY = [&](int jj) -> const double * { return sSY + (jj - (j0 - 1)) * 2 * RA; };
    const bool yside = threadIdx.x >= NC;
    const int col = yside ? (int)threadIdx.x - NC : (int)threadIdx.x;
    const int tx = col % TX, ty = col / TX;
    const int i = i0 + tx, j = j0 + ty;
    const bool active = (i < nxg) && (j < nyg);
    const bool wxl = (i == 0 || i == nxg - 1), wyl = (j == 0 || j == nyg - 1);
    const bool interior = active && !wxl && !wyl;
    double o[NZ];
    if (yside) {
        if (interior) {
            const int wsy = cclamp(j - A, 0, nyg - RA);
            double dy[NZ];
            const double *S1 = SY(j), *S2 = S1 + RA;
#pragma unroll
            for (int k = 0; k < NZ; k++) {
                double d1 = 0.0, d2 = 0.0;
#pragma unroll
                for (int m = 0; m < RA; m++) { const double v = P(k, i, wsy + m); d1 += S1[m] * v; d2 += S2[m] * v; }
                dy[k] = d1; o[k] = d2;
            }
            double dyc0 = dy[0];
            {
                int sh = 0;
                if (wsy == 0) sh = 1;
                if (wsy + RA - 1 == nyg - 1) sh = -1;
                if (sh != 0) {
                    const double *S = SY(j - sh);
                    double t = 0.0;
#pragma unroll
                    for (int m = 0; m < RA; m++) t += S[m] * P(0, i, wsy + sh + m);
                    dyc0 = t;
                }
            }
            const double *cfp = g.cf + i + (size_t)nxg * j;
            const double A2 = cfp[0], B2 = cfp[nxy], A4 = cfp[2 * (size_t)nxy], B4 = cfp[3 * (size_t)nxy];
            const double S5 = cfp[6 * (size_t)nxy];
            // the second sigma derivative is this side's share of the column's own terms (the x side takes
            // the first), so that both halves of a column do the same work
            double pc[NZ];
#pragma unroll
            for (int k = 0; k < NZ; k++) pc[k] = P(k, i, j);
#pragma unroll
            for (int k = 1; k < NZ - 1; k++) {
                const int wz = cclamp(k - A, 0, NZ - RA);
                const double *Z1 = g.sz + (k * 2) * RA, *Z2 = Z1 + RA;
                double dyz = 0.0, dzz = 0.0;
#pragma unroll
                for (int m = 0; m < RA; m++) {
                    dyz += Z1[m] * ((wz + m == 0) ? dyc0 : dy[wz + m]);
                    dzz += Z2[m] * pc[wz + m];
                }
                const double sg = g.zs[k];
                const double s2 = A2 + sg * B2, s4 = A4 + sg * B4;
                o[k] = o[k] + 2.0 * (s4 * dyz) + (s2 * s2 + s4 * s4 + S5 * S5) * dzz;
            }
            const double d = 1.0 / S5;
            o[0] = (A4 * d) * dy[1];   // h_y dy at the bottom row (h_y = A4 d, d = 1 / S5)
#pragma unroll
            for (int k = 0; k < NZ - 1; k++) sYp[k * NC + col] = o[k];
        }
    } else if (active) {
        if (wxl && wyl) {
#pragma unroll
            for (int k = 0; k < NZ; k++) o[k] = P(k, i, j);
        } else if (wxl || wyl) {
            const int ib = wxl ? ((i == 0) ? 1 : nxg - 2) : ((j == 0) ? 1 : nyg - 2);
            const int ws = wxl ? cclamp(ib - A, 0, nxg - RA) : cclamp(ib - A, 0, nyg - RA);
            const double *S = wxl ? SX(ib) : SY(ib);
            o[0] = P(0, i, j);
#pragma unroll
            for (int k = 1; k < NZ; k++) {
                double s = 0.0;
#pragma unroll
                for (int m = 0; m < RA; m++) s += S[m] * (wxl ? P(k, ws + m, j) : P(k, i, ws + m));
                o[k] = s;
            }
        } else {
            double dx[NZ], pc[NZ];
            const int wsx = cclamp(i - A, 0, nxg - RA);
            {
                const double *S1 = SX(i), *S2 = S1 + RA;
#pragma unroll
                for (int k = 0; k < NZ; k++) {
                    double d1 = 0.0, d2 = 0.0;
#pragma unroll
                    for (int m = 0; m < RA; m++) { const double v = P(k, wsx + m, j); d1 += S1[m] * v; d2 += S2[m] * v; }
                    dx[k] = d1; o[k] = d2;
                }
            }
#pragma unroll
            for (int k = 0; k < NZ; k++) pc[k] = P(k, i, j);
            double dxc0 = dx[0];
            {
                int sh = 0;
                if (wsx == 0) sh = 1;
                if (wsx + RA - 1 == nxg - 1) sh = -1;
                if (sh != 0) {
                    const double *S = SX(i - sh);
                    double t = 0.0;
#pragma unroll
                    for (int m = 0; m < RA; m++) t += S[m] * P(0, wsx + sh + m, j);
                    dxc0 = t;
                }
            }
            const double *cfp = g.cf + i + (size_t)nxg * j;
            const double A2 = cfp[0], B2 = cfp[nxy], A4 = cfp[2 * (size_t)nxy], B4 = cfp[3 * (size_t)nxy];
            const double A3 = cfp[4 * (size_t)nxy], B3 = cfp[5 * (size_t)nxy], S5 = cfp[6 * (size_t)nxy];
            {
                constexpr int wz = (1 - A) < 0 ? 0 : (1 - A);
                const double *Z1 = g.sz + (1 * 2) * RA;
                double dz = 0.0;
#pragma unroll
                for (int m = 0; m < RA; m++) dz += Z1[m] * pc[wz + m];
                const double sg = g.zs[1], d = 1.0 / S5, hx = A2 * d, hy = A4 * d;
                o[0] = (S5 + hx * (A2 + sg * B2) + hy * (A4 + sg * B4)) * dz + hx * dx[1];
            }
#pragma unroll
            for (int k = 1; k < NZ - 1; k++) {
                const int wz = cclamp(k - A, 0, NZ - RA);
                const double *Z1 = g.sz + (k * 2) * RA;
                double dz = 0.0, dxz = 0.0;
#pragma unroll
                for (int m = 0; m < RA; m++) {
                    const double z1 = Z1[m];
                    dz += z1 * pc[wz + m];
                    dxz += z1 * ((wz + m == 0) ? dxc0 : dx[wz + m]);
                }
                const double sg = g.zs[k];
                const double s2 = A2 + sg * B2, s3 = A3 + sg * B3;
                o[k] = o[k] + s3 * dz + 2.0 * (s2 * dxz);
            }
            o[NZ - 1] = pc[NZ - 1];
        }
    }
    __syncthreads();   // every read of the phi window and every y-side part done
    if (!yside && active) {
        if (interior) {
#pragma unroll
            for (int k = 0; k < NZ - 1; k++) o[k] += sYp[k * NC + col];
        }
#pragma unroll
        for (int k = 0; k < NZ; k++) sP[k * PS + (j - loy) * WX + (i - lox)] = o[k];
    }
    __syncthreads();
    {
        const int ie = cmin(i0 + TX, nxg), je = cmin(j0 + TY, nyg);
        for (int jj = j0; jj < je; jj++) {
            double *dst = out + ((size_t)jj * nxg + i0) * NZ;
            for (int e = threadIdx.x; e < (ie - i0) * NZ; e += NT) {
                const int ii = e / NZ, k = e - ii * NZ;
                dst[e] = sP[k * PS + (jj - loy) * WX + (i0 + ii - lox)];
            }
        }
    }
}

// x-z problems (Ny == 1), one thread per grid point: block = 64 columns x NZ levels.
template <int NZ, int A>
__global__ __launch_bounds__(256) void k_matvec2p(MvArgs g, const double *__restrict__ phi, double *__restrict__ out) {
    if (solve_done(g.skip)) return;
    constexpr int NT = 256, TX = 64, RA = 2 * A + 1, R = A + 1, WX = TX + 2 * R, NP = TX * NZ, PER = (NP + NT - 1) / NT;
    __shared__ double sP[WX * NZ];
    __shared__ double sDx[NP];
    const int nxg = g.nxg, nxy = g.nxy;
    const int i0 = blockIdx.x * TX;
    const int lox = cmax(cmin(i0 - R, nxg - 2 - 2 * A), 0);
    const int wx = cmin(WX, nxg - lox);
    stage_window<double, NZ, NZ, 1, WX, WX, NT>(sP, phi + (size_t)lox * NZ, 0, 1, wx);
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
// RA: the stencil width 2a + 1 when known at compile time (loads of all taps issued together), 0: g.a's.
template <int RA>
static __global__ __launch_bounds__(256) void k_surface_x(PlaneArgs g) {
    const int j = blockIdx.x, nxg = g.nxg, ra = RA > 0 ? RA : 2 * g.a + 1;
    double *E = g.E + (size_t)j * nxg, *P = g.P + (size_t)j * nxg;
    if (j >= g.gy && j < g.nyg - g.gy && threadIdx.x < 4) {
        double *v = (threadIdx.x & 1) ? P : E;
        if (threadIdx.x < 2) {
            const double *st = g.sx + (size_t)(1 * 2) * ra;
            double s = 0.0;
#pragma unroll
            for (int m = 1; m < (RA > 0 ? RA : ra); m++) s += st[m] * v[m];
            v[0] = (0.0 - s) / st[0];
        } else {
            const double *st = g.sx + (size_t)((nxg - 2) * 2) * ra;
            double s = 0.0;
#pragma unroll
            for (int m = 0; m < (RA > 0 ? RA : ra) - 1; m++) s += st[m] * v[nxg - ra + m];
            v[nxg - 1] = (0.0 - s) / st[ra - 1];
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nxg; i += blockDim.x) {
        const int ws = cclamp(i - g.a, 0, nxg - ra);
        const double *S1 = g.sx + (size_t)(i * 2) * ra, *S2 = S1 + ra;
        double e1 = 0.0, e2 = 0.0, p1 = 0.0;
#pragma unroll
        for (int m = 0; m < (RA > 0 ? RA : ra); m++) { e1 += S1[m] * E[ws + m]; e2 += S2[m] * E[ws + m]; p1 += S1[m] * P[ws + m]; }
        const size_t q = (size_t)j * nxg + i;
        g.Ex[q] = e1; g.Exx[q] = e2; g.Px[q] = p1;
        if (g.nyg == 1) {  // x-z problem: coefficient fields complete here
            const double d = g.nonlinear ? E[i] + g.h[q] : g.h[q];
            column_coefs(d, g.nonlinear ? e1 : 0.0, 0.0, g.nonlinear ? e2 : 0.0, g.bath[q], 0.0, g.bath[2 * g.nxy + q], g.cf, g.nxy, (int)q);
        }
    }
}

// y ghost values (columns with i in [gx, nxg-gx)), y derivatives and coefficients: one thread per plane
// point, 64 consecutive columns x 4 rows per block so that every load and store of a wave is contiguous.
// A ghost value depends only on interior values of its column, so each thread whose stencil reaches a
// ghost row evaluates it itself; the threads of the ghost rows also write it back.
template <int RB>
static __global__ __launch_bounds__(256) void k_surface_y(PlaneArgs g) {
    const int nxg = g.nxg, nyg = g.nyg, rb = RB > 0 ? RB : 2 * g.b + 1;
    const int i = (int)blockIdx.x * 64 + (int)(threadIdx.x & 63), j = (int)blockIdx.y * 4 + (int)(threadIdx.x >> 6);
    if (i >= nxg || j >= nyg) return;
    const bool gcol = i >= g.gx && i < nxg - g.gx;
    auto val = [&](const double *v, int jj) -> double {
        if (gcol && jj == 0) {
            const double *st = g.sy + (size_t)(1 * 2) * rb;
            double s = 0.0;
#pragma unroll
            for (int m = 1; m < (RB > 0 ? RB : rb); m++) s += st[m] * v[i + (size_t)nxg * m];
            return (0.0 - s) / st[0];
        }
        if (gcol && jj == nyg - 1) {
            const double *st = g.sy + (size_t)((nyg - 2) * 2) * rb;
            double s = 0.0;
#pragma unroll
            for (int m = 0; m < (RB > 0 ? RB : rb) - 1; m++) s += st[m] * v[i + (size_t)nxg * (nyg - rb + m)];
            return (0.0 - s) / st[rb - 1];
        }
        return v[i + (size_t)nxg * jj];
    };
    const int ws = cclamp(j - g.b, 0, nyg - rb);
    const double *S1 = g.sy + (size_t)(j * 2) * rb, *S2 = S1 + rb;
    double e1 = 0.0, e2 = 0.0, p1 = 0.0;
#pragma unroll
    for (int m = 0; m < (RB > 0 ? RB : rb); m++) {
        const double ev = val(g.E, ws + m), pv = val(g.P, ws + m);
        e1 += S1[m] * ev; e2 += S2[m] * ev; p1 += S1[m] * pv;
    }
    const size_t q = i + (size_t)nxg * j;
    const double eq = val(g.E, j);
    g.Ey[q] = e1; g.Eyy[q] = e2; g.Py[q] = p1;
    if (g.nonlinear)
        column_coefs(eq + g.h[q], g.Ex[q], e1, g.Exx[q] + e2, g.bath[q], g.bath[g.nxy + q], g.bath[2 * g.nxy + q], g.cf, g.nxy, (int)q);
    else
        column_coefs(g.h[q], 0.0, 0.0, 0.0, g.bath[q], g.bath[g.nxy + q], g.bath[2 * g.nxy + q], g.cf, g.nxy, (int)q);
    if (gcol && (j == 0 || j == nyg - 1)) {   // every reader above evaluated the ghost value itself
        const double pq = val(g.P, j);
        g.E[q] = eq; g.P[q] = pq;
    }
}

// W = dsigma/dz * one-sided sigma derivative at the surface (VerticalFreeSurfaceVelocity.f90:17).
// RC: the stencil width when known at compile time, 0: rc.
template <int RC>
static __global__ void k_vertical_velocity(int nxy, int nzg, int rc, const double *__restrict__ szTop,
                                    const double *__restrict__ cf, const double *__restrict__ phi, double *__restrict__ W) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nxy) return;
    if (RC > 0) rc = RC;
    const double *col = phi + (size_t)p * nzg + (nzg - rc);
    double s = 0.0;
#pragma unroll
    for (int m = 0; m < (RC > 0 ? RC : rc); m++) s += szTop[m] * col[m];
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
static __global__ void k_stage(StageArgs a) {
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

// Fluid velocities at every grid point from phi (StoreKinematicData.f90:266-326): W = dphi/dsigma, U =
// dphi/dx and V = dphi/dy with the path's first-derivative stencil rows (off-centred near the ends as
// DiffZArbitrary / DiffXEven / DiffYEven), in nonlinear mode plus the chain-rule terms
// ((1 - z)/d h_x - z/d eta_x) W and ((1 - z)/d h_y - z/d eta_y) W (z: the level's sigma coordinate, the
// ghost level's included; d = h + eta); the vertical velocity is W / d.  One thread per point.
struct KinArgs {
    int nxg, nyg, nzg, a, b, c, nonlinear;
    const double *sx, *sy, *sz;    // stencil rows [N][2][2a+1] (first derivative first)
    const double *z;               // sigma coordinate of each level, ghost level included
    const double *h, *bath;        // depth; bath planes h_x, h_y
    const double *E, *Ex, *Ey;     // surface and its derivatives
};
static __global__ __launch_bounds__(256) void k_kinematics(KinArgs a, const double *__restrict__ phi, double *__restrict__ U,
                                                           double *__restrict__ V, double *__restrict__ Wv,
                                                           double *__restrict__ Ws = nullptr) {
    const size_t nxy = (size_t)a.nxg * a.nyg, n = nxy * a.nzg;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(q % a.nzg);
        const size_t p = q / a.nzg;
        const int i = (int)(p % a.nxg), j = (int)(p / a.nxg);
        const int rz = 2 * a.c + 1, wz = cclamp(k - a.c, 0, a.nzg - rz);
        const double *Sz = a.sz + (size_t)(k * 2) * rz;
        double w = 0.0;
        for (int m = 0; m < rz; m++) w += Sz[m] * phi[p * a.nzg + wz + m];
        double u = 0.0, v = 0.0;
        if (a.nxg > 1) {
            const int rx = 2 * a.a + 1, wx = cclamp(i - a.a, 0, a.nxg - rx);
            const double *Sx = a.sx + (size_t)(i * 2) * rx;
            for (int m = 0; m < rx; m++) u += Sx[m] * phi[((size_t)j * a.nxg + wx + m) * a.nzg + k];
        }
        if (a.nyg > 1) {
            const int ry = 2 * a.b + 1, wy = cclamp(j - a.b, 0, a.nyg - ry);
            const double *Sy = a.sy + (size_t)(j * 2) * ry;
            for (int m = 0; m < ry; m++) v += Sy[m] * phi[((size_t)(wy + m) * a.nxg + i) * a.nzg + k];
        }
        const double d = a.h[p] + a.E[p], z = a.z[k];
        if (a.nonlinear) {
            if (a.nxg > 1) u = u + ((1 - z) / d * a.bath[p] - z / d * a.Ex[p]) * w;
            if (a.nyg > 1) v = v + ((1 - z) / d * a.bath[nxy + p] - z / d * a.Ey[p]) * w;
        }
        U[q] = u; V[q] = v; Wv[q] = w / d;
        if (Ws) Ws[q] = w;
    }
}

// The velocity gradients StoreKinematicData.f90:327-478 writes after u, v, w, from F = (W, U, V) with
// W = dphi/dsigma (not divided by d) and U, V the velocities above: for each field, G = Dsigma(F) / d,
// Dx(F) + cx Dsigma(F), Dy(F) + cy Dsigma(F) with the chain-rule factors cx = (1 - z)/d h_x - z/d eta_x,
// cy likewise (nonlinear mode only), x (y) derivatives zero when the grid has one column (row).  G holds
// the 9 fields of n points in that order (for W: d w_z, (d w)_x, (d w)_y).  One thread per point.
static __global__ __launch_bounds__(256) void k_kinematics_grad(KinArgs a, const double *__restrict__ W,
                                                                const double *__restrict__ U, const double *__restrict__ V,
                                                                double *__restrict__ G) {
    const size_t nxy = (size_t)a.nxg * a.nyg, n = nxy * a.nzg;
    const double *F[3] = {W, U, V};
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(q % a.nzg);
        const size_t p = q / a.nzg;
        const int i = (int)(p % a.nxg), j = (int)(p / a.nxg);
        const int rz = 2 * a.c + 1, wz = cclamp(k - a.c, 0, a.nzg - rz);
        const double *Sz = a.sz + (size_t)(k * 2) * rz;
        const double d = a.h[p] + a.E[p], z = a.z[k];
        const double cx = ((1 - z) / d * a.bath[p] - z / d * a.Ex[p]);
        const double cy = ((1 - z) / d * a.bath[nxy + p] - z / d * a.Ey[p]);
        for (int f = 0; f < 3; f++) {
            const double *Fv = F[f];
            double fz = 0.0;
            for (int m = 0; m < rz; m++) fz += Sz[m] * Fv[p * a.nzg + wz + m];
            double fx = 0.0, fy = 0.0;
            if (a.nxg > 1) {
                const int rx = 2 * a.a + 1, wx = cclamp(i - a.a, 0, a.nxg - rx);
                const double *Sx = a.sx + (size_t)(i * 2) * rx;
                for (int m = 0; m < rx; m++) fx += Sx[m] * Fv[((size_t)j * a.nxg + wx + m) * a.nzg + k];
                if (a.nonlinear) fx = fx + cx * fz;
            }
            if (a.nyg > 1) {
                const int ry = 2 * a.b + 1, wy = cclamp(j - a.b, 0, a.nyg - ry);
                const double *Sy = a.sy + (size_t)(j * 2) * ry;
                for (int m = 0; m < ry; m++) fy += Sy[m] * Fv[((size_t)(wy + m) * a.nxg + i) * a.nzg + k];
                if (a.nonlinear) fy = fy + cy * fz;
            }
            G[(size_t)(3 * f) * n + q] = fz / d;
            G[(size_t)(3 * f + 1) * n + q] = fx;
            G[(size_t)(3 * f + 2) * n + q] = fy;
        }
    }
}

// Laplace right-hand side: Dirichlet data P on the free surface of non-ghost columns, zero elsewhere
// (Runge_Kutta_4.f90:65-69 with no wall flux).
static __global__ void k_fill_rhs(int nxg, int nyg, int nzg, int gx, int gy, const double *__restrict__ P, double *__restrict__ b) {
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
static __global__ void k_axpby(size_t n, double a, const double *__restrict__ x, double b, double *__restrict__ y,
                               const int *skip = nullptr) {
    if (solve_done(skip)) return;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        y[q] = a * x[q] + b * y[q];
}
// y = b - y
static __global__ void k_bminus(size_t n, const double *__restrict__ b, double *__restrict__ y, const int *skip) {
    if (solve_done(skip)) return;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        y[q] = b[q] - y[q];
}
template <typename S, typename D>
__global__ void k_convert(size_t n, const S *__restrict__ x, D *__restrict__ y, const int *skip) {
    if (solve_done(skip)) return;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        y[q] = (D)x[q];
}
// Preconditioner input: b = (T) r and the iterate x = 0 in one pass.
template <typename T>
__global__ void k_precond_in(size_t n, const double *__restrict__ r, T *__restrict__ b, T *__restrict__ x, const int *skip) {
    if (solve_done(skip)) return;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        b[q] = (T)r[q];
        x[q] = T(0);
    }
}
template <typename T>
__global__ void k_fill(size_t n, T v, T *__restrict__ y, const int *skip) {
    if (solve_done(skip)) return;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        y[q] = v;
}

// Block reduction of MAXK accumulators into part[blockIdx.x*MAXK + i].
// Block sum of NA per-thread accumulators, stored at part[blockIdx.x * PSTRIDE + i] (fixed order).
template <int NA, int PSTRIDE>
__device__ __forceinline__ void block_reduce_store(const double (&acc)[NA], double *part) {
    __shared__ double red[NA][4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NA; i++) {
        double v = acc[i];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) red[i][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < NA) {
        const int i = threadIdx.x;
        part[(size_t)blockIdx.x * PSTRIDE + i] = ((red[i][0] + red[i][1]) + red[i][2]) + red[i][3];
    }
}

// The Krylov kernels below take the number of basis vectors K as a template parameter: every load of a
// grid point is issued before the first use (no predicated loads, no waits between them).

// part[b][i] = sum over block b's range of V_i . w for i < K (V_i = V + i*n); NORM adds w . w at index K.
template <int K, int NORM, int PSTRIDE>
__global__ __launch_bounds__(256) void k_mdot(size_t n, const double *__restrict__ V, const double *__restrict__ w,
                                               double *__restrict__ part, const int *skip) {
    if (solve_done(skip)) return;
    double acc[K + NORM];
#pragma unroll
    for (int i = 0; i < K + NORM; i++) acc[i] = 0.0;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        double v[K];
#pragma unroll
        for (int i = 0; i < K; i++) v[i] = V[(size_t)i * n + q];
        const double wq = w[q];
#pragma unroll
        for (int i = 0; i < K; i++) acc[i] += v[i] * wq;
        if (NORM) acc[K + NORM - 1] += wq * wq;
    }
    block_reduce_store<K + NORM, PSTRIDE>(acc, part);
}

// v <- v - sum_i h_i V_i; part gets the dots of the new v with V_i (i < K) and its squared norm (index K).
template <int K, int PSTRIDE>
__global__ __launch_bounds__(256) void k_gs_update_dot(size_t n, const double *__restrict__ V, const double *__restrict__ h,
                                                        double *__restrict__ v, double *__restrict__ part, const int *skip) {
    if (solve_done(skip)) return;
    double acc[K + 1], hh[K];
#pragma unroll
    for (int i = 0; i < K; i++) { acc[i] = 0.0; hh[i] = h[i]; }
    acc[K] = 0.0;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        double vv[K];
#pragma unroll
        for (int i = 0; i < K; i++) vv[i] = V[(size_t)i * n + q];
        double x = v[q];
#pragma unroll
        for (int i = 0; i < K; i++) x -= hh[i] * vv[i];
        v[q] = x;
#pragma unroll
        for (int i = 0; i < K; i++) acc[i] += vv[i] * x;
        acc[K] += x * x;
    }
    block_reduce_store<K + 1, PSTRIDE>(acc, part);
}

// v <- (v + sign * sum_i h_i V_i) * scale  (sign -1: orthogonalisation; +1, scale 1: solution update);
// scale_dev, when given, holds the scale in device memory.
template <int K>
__global__ __launch_bounds__(256) void k_gs_update(size_t n, const double *__restrict__ V, const double *__restrict__ h,
                                                    double sign, double scale, const double *scale_dev, double *__restrict__ v,
                                                    const int *skip) {
    if (solve_done(skip)) return;
    if (scale_dev) scale = *scale_dev;
    double hh[K];
#pragma unroll
    for (int i = 0; i < K; i++) hh[i] = sign * h[i];
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        double vv[K];
#pragma unroll
        for (int i = 0; i < K; i++) vv[i] = V[(size_t)i * n + q];
        double x = v[q];
#pragma unroll
        for (int i = 0; i < K; i++) x += hh[i] * vv[i];
        v[q] = x * scale;
    }
}

// out[i] = sum_b part[b][i]; one block per entry, fixed summation order (deterministic).
template <int MAXK>
__global__ __launch_bounds__(256) void k_reduce_parts(int nblk, int k, const double *part, double *out, const int *skip) {
    if (solve_done(skip)) return;
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
// Device-driven Krylov solves: the scalar work of GMRES(m) (Hessenberg column, Givens rotations,
// stopping test, back substitution) and of defect correction runs in one-thread kernels on the solve's
// stream, so a whole solve is a fixed launch sequence with no host round trip.  After convergence the
// solve's kernels see `done` and return at once.
// ---------------------------------------------------------------------------------------------
constexpr int KRYLOV_MAXK = 32;   // basis limit: GMRESmaxiterations <= KRYLOV_MAXK - 1

struct KrylovState {
    int done;       // solve finished (converged, breakdown or out of iterations)
    int skip2;      // 0 only between the two passes of an Arnoldi step that needs the DGKS second pass
    int k;          // Arnoldi steps taken
    int kk;         // basis vectors combined into the solution
    int matvecs;    // operator applications of this solve
    int status;     // 0 converged, 1 not converged
    double beta, tol, est, scale;
    double H[KRYLOV_MAXK * KRYLOV_MAXK];   // Hessenberg entry (i, j) at H[i * KRYLOV_MAXK + j]
    double cs[KRYLOV_MAXK], sn[KRYLOV_MAXK], g[KRYLOV_MAXK + 1], h1[KRYLOV_MAXK], y[KRYLOV_MAXK];
    // over all solves of the handle
    long long it_total, it_solves;
    int it_min, it_max, status_any;
    // speculative RK4 steps: a solve enqueued with fewer iteration slots than its budget that used them all
    // without stopping sets overflow; amax[stage] = most slots a solve of that stage used in this step
    int overflow, amax[4];
    long long sv_total, sv_solves;   // statistics at the start of the step (restored when it is replayed)
    int sv_min, sv_max, sv_status;
    // one-reduction CGS2 (left preconditioning): the Hessenberg columns before rotation, and the pending
    // basis vector's second-pass projections cv and inverse norm ib (the vector is normalised in the next
    // step's first pass)
    double Hr[KRYLOV_MAXK * KRYLOV_MAXK];
    double cv[KRYLOV_MAXK], ib;
};

// Start of a speculative RK4 step: clear the per-step report, keep the statistics for a replay.
static __global__ void k_step_begin(KrylovState *S) {
    if (threadIdx.x != 0) return;
    S->overflow = 0;
    for (int i = 0; i < 4; i++) S->amax[i] = 0;
    S->sv_total = S->it_total; S->sv_solves = S->it_solves; S->sv_min = S->it_min; S->sv_max = S->it_max;
    S->sv_status = S->status_any;
}
// Replay of a step: statistics as they were before it.
static __global__ void k_step_restore(KrylovState *S) {
    if (threadIdx.x != 0) return;
    S->it_total = S->sv_total; S->it_solves = S->sv_solves; S->it_min = S->sv_min; S->it_max = S->sv_max;
    S->status_any = S->sv_status;
}
// End of a step: overflow flag and per-stage slot use to host-visible memory.
static __global__ void k_step_report(const KrylovState *S, int *out) {
    if (threadIdx.x != 0) return;
    out[0] = S->overflow;
    for (int i = 0; i < 4; i++) out[1 + i] = S->amax[i];
}

// Start of a solve (before its first operator application).
static __global__ void k_solve_begin(KrylovState *S) {
    if (threadIdx.x != 0) return;
    S->done = 0; S->skip2 = 1; S->k = 0; S->kk = 0; S->matvecs = 1; S->status = 1;
}

// GMRES: hv[0] = |M^-1 (b - A x0)|^2.  Sets the stopping tolerance (rtol * beta + atol) and the scale of
// the first basis vector.
static __global__ void k_gmres_start(KrylovState *S, const double *hv, double rtol, double atol) {
    if (threadIdx.x != 0) return;
    const double beta = sqrt(hv[0]);
    S->beta = beta; S->tol = rtol * beta + atol; S->est = beta;
    if (beta <= S->tol) { S->done = 1; S->status = 0; return; }
    S->scale = 1.0 / beta;
    S->ib = 1.0;
    S->g[0] = beta;
}

// Hessenberg column k (1-based step): col[0..k-1] projections, col[k] = hn the new vector's norm, held in
// LDS with the previous rotations (cs, sn).  One lane applies the previous rotations and the new one and
// updates the residual estimate and the stopping test; the caller writes col back as column k - 1 of H.
__device__ inline void gmres_column(KrylovState *S, int k, double *col, const double *cs, const double *sn, int m) {
    constexpr int M = KRYLOV_MAXK;
    const double hn = col[k];
    for (int i = 0; i < k - 1; i++) {
        const double a = col[i], b = col[i + 1];
        col[i] = cs[i] * a + sn[i] * b;
        col[i + 1] = cs[i] * b - sn[i] * a;
    }
    const double a = col[k - 1], b = col[k];
    const double rr = sqrt(a * a + b * b);
    double c = 1.0, sv = 0.0;
    if (rr > 0.0) { c = a / rr; sv = b / rr; }
    S->cs[k - 1] = c; S->sn[k - 1] = sv;
    col[k - 1] = rr;
    col[k] = hn;
    const double g0 = S->g[k - 1];
    S->g[k] = -sv * g0;
    S->g[k - 1] = c * g0;
    const double est = fabs(-sv * g0);
    S->est = est; S->k = k;
    S->scale = hn > 0.0 ? 1.0 / hn : 1.0;
    if (!(k < m && !(est <= S->tol) && S->matvecs < m) || hn == 0.0 || k + 1 >= M) S->done = 1;
}

// Arnoldi step k, first pass: hv[0..k-1] = V_i . w, hv[k] = w . w (classical Gram-Schmidt).  Keeps the
// column when the projection removed at most half of the squared norm, else asks for the second pass.
// One wave: the vectors are staged in LDS by all lanes, the serial part runs on lane 0.
static __global__ __launch_bounds__(64) void k_gmres_pass1(KrylovState *S, const double *hv, int k, int m) {
    constexpr int M = KRYLOV_MAXK;
    __shared__ double col[M + 1], cs[M], sn[M];
    __shared__ int second;
    if (S->done) return;
    const int t = threadIdx.x;
    if (t <= k) col[t] = hv[t];
    if (t < k - 1) { cs[t] = S->cs[t]; sn[t] = S->sn[t]; }
    __syncthreads();
    if (t == 0) {
        S->matvecs += 1;
        const double nrm2 = col[k];
        double nn = nrm2;
        for (int i = 0; i < k; i++) nn -= col[i] * col[i];
        second = !(nn > 0.5 * nrm2);
        if (!second) {
            col[k] = sqrt(nn);
            gmres_column(S, k, col, cs, sn, m);
        }
        S->skip2 = second ? 0 : 1;
    }
    __syncthreads();
    if (second) {
        if (t < k) S->h1[t] = col[t];
    } else if (t <= k) {
        S->H[t * M + (k - 1)] = col[t];
    }
}

// Arnoldi step k, second (DGKS) pass: hv holds the projections of the once-orthogonalised vector.
static __global__ __launch_bounds__(64) void k_gmres_pass2(KrylovState *S, const double *hv, int k, int m) {
    constexpr int M = KRYLOV_MAXK;
    __shared__ double col[M + 1], h2[M + 1], cs[M], sn[M];
    if (S->skip2) return;
    const int t = threadIdx.x;
    if (t <= k) h2[t] = hv[t];
    if (t < k) col[t] = S->h1[t];
    if (t < k - 1) { cs[t] = S->cs[t]; sn[t] = S->sn[t]; }
    __syncthreads();
    if (t == 0) {
        double n2 = h2[k];
        for (int i = 0; i < k; i++) { n2 -= h2[i] * h2[i]; col[i] = col[i] + h2[i]; }
        col[k] = sqrt(fmax(n2, 0.0));
        gmres_column(S, k, col, cs, sn, m);
        S->skip2 = 1;
    }
    __syncthreads();
    if (t <= k) S->H[t * M + (k - 1)] = col[t];
}

// One-reduction classical Gram-Schmidt with reorthogonalisation (DCGS2: Swirydowicz, Langou, Ananthan,
// Yang and Thomas, Numer. Linear Algebra Appl. 28 (2021) e2343; Bielich et al., Parallel Comput. 112
// (2022) 102940) for left-preconditioned GMRES.  Step k (1-based) applies the operator to the candidate u
// in slot k-1 -- the once-orthogonalised, unnormalised vector of step k-1 -- rather than to the basis
// vector V_{k-1} = (u - sum_{i<k-1} c_i V_i) ib, whose second-pass projections c and inverse norm ib were
// found in step k-1.  By linearity z = M^-1 A u = (1/ib) M^-1 A V_{k-1} + V H c (H the unrotated
// Hessenberg columns 0..k-2), so with d = V^T z over V_0..V_{k-1}:
//   first pass: V_{k-1} normalised in place and d in one read of the basis (k_dcgs_a),
//   second pass: u' = (z - V d) ib and its projections c' = V^T u', |u'|^2 in one more read (k_dcgs_b),
//   column k-1: h = (d - H c) ib, H(0..k-1, k-1) = h + c', H(k, k-1) = sqrt(|u'|^2 - |c'|^2).
// Two reads of the basis per step instead of three (projection, second projection, update).

// Step k >= 2, first pass: slot K-1 holds u; V_{K-1} = (u - sum_{i<K-1} c_i V_i) ib is written back, and
// part gets the dots of z (slot K) with V_0 .. V_{K-1}.
template <int K, int PSTRIDE>
__global__ __launch_bounds__(256) void k_dcgs_a(size_t n, double *__restrict__ V, const KrylovState *S,
                                                 double *__restrict__ part, const int *skip) {
    static_assert(K >= 2, "the first step has no pending vector");
    if (solve_done(skip)) return;
    double c[K - 1], acc[K];
#pragma unroll
    for (int i = 0; i < K - 1; i++) c[i] = S->cv[i];
    const double ib = S->ib;
#pragma unroll
    for (int i = 0; i < K; i++) acc[i] = 0.0;
    double *__restrict__ u = V + (size_t)(K - 1) * n;
    const double *__restrict__ z = V + (size_t)K * n;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        double vv[K - 1];
#pragma unroll
        for (int i = 0; i < K - 1; i++) vv[i] = V[(size_t)i * n + q];
        double x = u[q];
        const double zq = z[q];
#pragma unroll
        for (int i = 0; i < K - 1; i++) x -= c[i] * vv[i];
        x *= ib;
        u[q] = x;
#pragma unroll
        for (int i = 0; i < K - 1; i++) acc[i] += vv[i] * zq;
        acc[K - 1] += x * zq;
    }
    block_reduce_store<K, PSTRIDE>(acc, part);
}

// Step k, second pass: v <- (v - sum_{i<K} h_i V_i) ib (h the first pass's dots); part gets the dots of the
// new v with V_i (i < K) and its squared norm (index K).
template <int K, int PSTRIDE>
__global__ __launch_bounds__(256) void k_dcgs_b(size_t n, const double *__restrict__ V, const double *__restrict__ h,
                                                 const KrylovState *S, double *__restrict__ v, double *__restrict__ part,
                                                 const int *skip) {
    if (solve_done(skip)) return;
    double acc[K + 1], hh[K];
#pragma unroll
    for (int i = 0; i < K; i++) { acc[i] = 0.0; hh[i] = h[i]; }
    acc[K] = 0.0;
    const double ib = S->ib;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        double vv[K];
#pragma unroll
        for (int i = 0; i < K; i++) vv[i] = V[(size_t)i * n + q];
        double x = v[q];
#pragma unroll
        for (int i = 0; i < K; i++) x -= hh[i] * vv[i];
        x *= ib;
        v[q] = x;
#pragma unroll
        for (int i = 0; i < K; i++) acc[i] += vv[i] * x;
        acc[K] += x * x;
    }
    block_reduce_store<K + 1, PSTRIDE>(acc, part);
}

// The two passes on pairs of grid points (16-byte loads and stores), for an even n: the basis vectors then
// start 16-byte aligned.  Each lane sums its pair's products in order (first point, then second).
template <int K, int PSTRIDE>
__global__ __launch_bounds__(256) void k_dcgs_a2(size_t n, double *__restrict__ V, const KrylovState *S,
                                                  double *__restrict__ part, const int *skip) {
    static_assert(K >= 2, "the first step has no pending vector");
    if (solve_done(skip)) return;
    double c[K - 1], acc[K];
#pragma unroll
    for (int i = 0; i < K - 1; i++) c[i] = S->cv[i];
    const double ib = S->ib;
#pragma unroll
    for (int i = 0; i < K; i++) acc[i] = 0.0;
    double2 *__restrict__ u = reinterpret_cast<double2 *>(V + (size_t)(K - 1) * n);
    const double2 *__restrict__ z = reinterpret_cast<const double2 *>(V + (size_t)K * n);
    const size_t n2 = n / 2;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n2; q += (size_t)gridDim.x * blockDim.x) {
        double2 vv[K - 1];
#pragma unroll
        for (int i = 0; i < K - 1; i++) vv[i] = reinterpret_cast<const double2 *>(V + (size_t)i * n)[q];
        double2 x = u[q];
        const double2 zq = z[q];
#pragma unroll
        for (int i = 0; i < K - 1; i++) { x.x -= c[i] * vv[i].x; x.y -= c[i] * vv[i].y; }
        x.x *= ib; x.y *= ib;
        u[q] = x;
#pragma unroll
        for (int i = 0; i < K - 1; i++) { acc[i] += vv[i].x * zq.x; acc[i] += vv[i].y * zq.y; }
        acc[K - 1] += x.x * zq.x; acc[K - 1] += x.y * zq.y;
    }
    block_reduce_store<K, PSTRIDE>(acc, part);
}
template <int K, int PSTRIDE>
__global__ __launch_bounds__(256) void k_dcgs_b2(size_t n, const double *__restrict__ V, const double *__restrict__ h,
                                                  const KrylovState *S, double *__restrict__ v, double *__restrict__ part,
                                                  const int *skip) {
    if (solve_done(skip)) return;
    double acc[K + 1], hh[K];
#pragma unroll
    for (int i = 0; i < K; i++) { acc[i] = 0.0; hh[i] = h[i]; }
    acc[K] = 0.0;
    const double ib = S->ib;
    double2 *__restrict__ v2 = reinterpret_cast<double2 *>(v);
    const size_t n2 = n / 2;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n2; q += (size_t)gridDim.x * blockDim.x) {
        double2 vv[K];
#pragma unroll
        for (int i = 0; i < K; i++) vv[i] = reinterpret_cast<const double2 *>(V + (size_t)i * n)[q];
        double2 x = v2[q];
#pragma unroll
        for (int i = 0; i < K; i++) { x.x -= hh[i] * vv[i].x; x.y -= hh[i] * vv[i].y; }
        x.x *= ib; x.y *= ib;
        v2[q] = x;
#pragma unroll
        for (int i = 0; i < K; i++) { acc[i] += vv[i].x * x.x; acc[i] += vv[i].y * x.y; }
        acc[K] += x.x * x.x; acc[K] += x.y * x.y;
    }
    block_reduce_store<K + 1, PSTRIDE>(acc, part);
}

// Step k: d = the first pass's dots (k), e = the second pass's dots and squared norm (k + 1).  Completes
// Hessenberg column k - 1, rotates it, updates the residual estimate and the stopping test, and keeps c'
// and the new inverse norm for the next step.
static __global__ __launch_bounds__(64) void k_dcgs_column(KrylovState *S, const double *d, const double *e, int k, int m) {
    constexpr int M = KRYLOV_MAXK;
    __shared__ double col[M + 1], cs[M], sn[M];
    if (S->done) return;
    const int t = threadIdx.x;
    if (t < k) {   // h_t = (d_t - sum_j H(t, j) c_j) ib over the columns j = t-1 .. k-2 that have row t
        double acc = d[t];
        for (int j = t > 0 ? t - 1 : 0; j <= k - 2; j++) acc -= S->Hr[t * M + j] * S->cv[j];
        col[t] = acc * S->ib + e[t];
    }
    if (t < k - 1) { cs[t] = S->cs[t]; sn[t] = S->sn[t]; }
    __syncthreads();
    if (t == 0) {
        S->matvecs += 1;
        double n2 = e[k];
        for (int i = 0; i < k; i++) n2 -= e[i] * e[i];
        col[k] = sqrt(fmax(n2, 0.0));
        for (int i = 0; i <= k; i++) S->Hr[i * M + (k - 1)] = col[i];
        for (int i = 0; i < k; i++) S->cv[i] = e[i];
        S->ib = col[k] > 0.0 ? 1.0 / col[k] : 1.0;
        gmres_column(S, k, col, cs, sn, m);
    }
    __syncthreads();
    if (t <= k) S->H[t * M + (k - 1)] = col[t];
}

__device__ inline void solve_stats(KrylovState *S) {
    S->it_total += S->matvecs; S->it_solves += 1;
    if (S->it_min < 0 || S->matvecs < S->it_min) S->it_min = S->matvecs;
    if (S->matvecs > S->it_max) S->it_max = S->matvecs;
    S->status_any |= S->status;
}

// End of a GMRES solve: back substitution H y = g on the kk leading columns (trailing zero pivots
// dropped), y to S->y; the solution update follows in k_basis_combine.  slots < full: the solve was
// enqueued with fewer Arnoldi steps than its budget (flags overflow when it used them all).
// (One wave stages the leading rows of H and g in LDS; lane 0 then substitutes.)
static __global__ __launch_bounds__(64) void k_gmres_finish(KrylovState *S, int slots, int full, int stage) {
    constexpr int M = KRYLOV_MAXK;
    __shared__ double sH[M * M], sg[M], sy[M];
    const int t = threadIdx.x, k = S->k;
    const bool solve = !(S->beta <= S->tol);
    if (solve) {
        for (int e = t; e < k * M; e += 64) sH[e] = S->H[e];
        if (t < k) sg[t] = S->g[t];
    }
    __syncthreads();
    if (t != 0) return;
    if (!S->done && slots < full) S->overflow = 1;
    if (k > S->amax[stage]) S->amax[stage] = k;
    int kk = 0;
    if (solve) {
        kk = k;
        while (kk > 0 && sH[(kk - 1) * M + (kk - 1)] == 0.0) kk--;
        for (int i = kk - 1; i >= 0; i--) {
            double acc = sg[i];
            for (int q = i + 1; q < kk; q++) acc -= sH[i * M + q] * sy[q];
            sy[i] = acc / sH[i * M + i];
            S->y[i] = sy[i];
        }
        S->status = (S->est <= S->tol) ? 0 : 1;
    }
    S->kk = kk;
    solve_stats(S);
}

// x += sum_{i < S->kk} S->y[i] V_i
static __global__ __launch_bounds__(256) void k_basis_combine(size_t n, const double *__restrict__ V, const KrylovState *S,
                                                              double *__restrict__ x) {
    const int kk = S->kk;
    if (kk == 0) return;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        double acc = x[q];
        for (int i0 = 0; i0 < kk; i0 += 8) {   // eight basis loads in flight per thread
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) v[u] = (i0 + u < kk) ? V[(size_t)(i0 + u) * n + q] : 0.0;
#pragma unroll
            for (int u = 0; u < 8; u++)
                if (i0 + u < kk) acc += S->y[i0 + u] * v[u];
        }
        x[q] = acc;
    }
}

// v *= S->scale
static __global__ void k_scale_state(size_t n, const KrylovState *S, double *__restrict__ v) {
    if (S->done) return;
    const double a = S->scale;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) v[q] *= a;
}

// Defect correction, iteration it (1-based): hv[0] = |z|^2 of the preconditioned defect z; the tolerance
// is set from the first one (dc.f:77-101).
static __global__ void k_dc_check(KrylovState *S, const double *hv, int it, double rtol, double atol) {
    if (threadIdx.x != 0 || S->done) return;
    const double alpha = sqrt(hv[0]);
    if (it == 1) { S->beta = alpha; S->tol = rtol * alpha + atol; }
    S->est = alpha; S->matvecs = it;
    if (alpha <= S->tol) { S->done = 1; S->status = 0; }
}

static __global__ void k_dc_finish(KrylovState *S, int slots, int full, int stage) {
    if (threadIdx.x != 0) return;
    if (!S->done && slots < full) S->overflow = 1;
    if (S->matvecs > S->amax[stage]) S->amax[stage] = S->matvecs;
    S->status = S->done ? S->status : 1;
    S->kk = 0;
    solve_stats(S);
}

// ---------------------------------------------------------------------------------------------
// Multigrid (low-order linearised operator, matrix-free; BuildLinearSystemMatrix.f90 row semantics)
// ---------------------------------------------------------------------------------------------
template <typename T>
struct LvArgs {
    int nxg, nyg, nzg, nxy, gx, gy, dim3;
    const T *sx, *sy, *sz;  // 3-point stencil rows [N][2][3]
    const T *zs;            // sigma of each level (0 on the ghost level)
    const T *col;           // 6 planes: 1/h, h_x, h_y, h_x/h, h_y/h, -2(h_x^2+h_y^2)/h^2 + (h_xx+h_yy)/h
    const int *skip;        // solve_done flag (nullptr: always run)
    int spec;               // 1: the kernels issue their loads before testing skip (small levels, where a
                            // launch is a chain of memory latencies); 0: skip is tested first
    // uniform horizontal grid (uni = 1): every interior column's centred 3-point x (y) rows are the same,
    // first derivative (-cx1, 0, cx1), second derivative (cx2, cxc, cx2)
    int uni;
    T cx1, cx2, cxc, cy1, cy2, cyc;
};

// Level coefficients as the column kernels read them: 3-point stencil rows sx(i), sy(j), sz(k) (first then
// second derivative, 6 values), sigma levels zs(k), and the per-column planes col(plane, i, j).
// CoefG reads the level arrays in HBM; CoefS reads copies staged in LDS for a window of columns.
template <typename T>
struct CoefG {
    const LvArgs<T> &L;
    __device__ __forceinline__ const T *sx(int i) const { return L.sx + i * 6; }
    __device__ __forceinline__ const T *sy(int j) const { return L.sy + j * 6; }
    __device__ __forceinline__ const T *sz(int k) const { return L.sz + k * 6; }
    __device__ __forceinline__ T zs(int k) const { return L.zs[k]; }
    __device__ __forceinline__ T col(int pl, int i, int j) const { return L.col[(size_t)pl * L.nxy + i + (size_t)L.nxg * j]; }
};
template <typename T>
struct CoefS {
    const T *sx_, *sy_, *sz_, *zs_, *col_;
    int ib, jb, nx, np;   // window origin, row length and plane size of the staged column planes
    __device__ __forceinline__ const T *sx(int i) const { return sx_ + (i - ib) * 6; }
    __device__ __forceinline__ const T *sy(int j) const { return sy_ + (j - jb) * 6; }
    __device__ __forceinline__ const T *sz(int k) const { return sz_ + k * 6; }
    __device__ __forceinline__ T zs(int k) const { return zs_[k]; }
    __device__ __forceinline__ T col(int pl, int i, int j) const { return col_[pl * np + (j - jb) * nx + (i - ib)]; }
};

// LDS sizes (elements) of the staged coefficients of a window of MW x MH columns and NZ levels.
template <int MW, int MH, int NZ>
struct CoefLds {
    static constexpr int SX = MW * 6, SY = MH * 6, SZ = NZ * 6, ZS = NZ, COL = 6 * MW * MH;
    static constexpr int TOTAL = SX + SY + SZ + ZS + COL;
};

// Stages the coefficients of the window [lox, lox+wcols) x [loy, loy+wrows) (row length MW) into buf
// (CoefLds<MW, MH, NZ>::TOTAL elements) and returns the accessor.  Each segment is one predicated load per
// element with no branches around it, so all loads of a thread are in flight together.
// CoefRegs holds the two halves separately (load into registers, store to LDS), as WindowRegs.
template <typename T, int NZ, int MW, int MH, int NT>
struct CoefRegs {
    using S = CoefLds<MW, MH, NZ>;
    static constexpr int PX = (S::SX + NT - 1) / NT, PY = (S::SY + NT - 1) / NT, PZ = (S::SZ + S::ZS + NT - 1) / NT;
    static constexpr int PC = (S::COL + NT - 1) / NT;
    T vx[PX], vy[PY], vz[PZ], vc[PC];
    int lox, loy;
    __device__ __forceinline__ void load(const LvArgs<T> &L, int lox_, int loy_, int wcols, int wrows) {
        lox = lox_; loy = loy_;
        const int t = (int)threadIdx.x;
#pragma unroll
        for (int q = 0; q < PX; q++) {
            const int e = t + q * NT;
            vx[q] = (e < S::SX && e / 6 < wcols) ? L.sx[lox * 6 + e] : T(0);
        }
#pragma unroll
        for (int q = 0; q < PY; q++) {
            const int e = t + q * NT;
            vy[q] = (e < S::SY && e / 6 < wrows) ? L.sy[loy * 6 + e] : T(0);
        }
#pragma unroll
        for (int q = 0; q < PZ; q++) {
            const int e = t + q * NT;
            vz[q] = e < S::SZ ? L.sz[e] : (e < S::SZ + S::ZS ? L.zs[e - S::SZ] : T(0));
        }
#pragma unroll
        for (int q = 0; q < PC; q++) {
            const int e = t + q * NT;
            const int pl = e / (MW * MH), c = e - pl * (MW * MH), r = c / MW, cc = c - r * MW;
            vc[q] = (e < S::COL && r < wrows && cc < wcols) ? L.col[(size_t)pl * L.nxy + (size_t)(loy + r) * L.nxg + lox + cc] : T(0);
        }
    }
    __device__ __forceinline__ CoefS<T> store(T *buf) const {
        const int t = (int)threadIdx.x;
#pragma unroll
        for (int q = 0; q < PX; q++) { const int e = t + q * NT; if (e < S::SX) buf[e] = vx[q]; }
#pragma unroll
        for (int q = 0; q < PY; q++) { const int e = t + q * NT; if (e < S::SY) buf[S::SX + e] = vy[q]; }
#pragma unroll
        for (int q = 0; q < PZ; q++) { const int e = t + q * NT; if (e < S::SZ + S::ZS) buf[S::SX + S::SY + e] = vz[q]; }
#pragma unroll
        for (int q = 0; q < PC; q++) { const int e = t + q * NT; if (e < S::COL) buf[S::SX + S::SY + S::SZ + S::ZS + e] = vc[q]; }
        CoefS<T> C;
        C.sx_ = buf; C.sy_ = buf + S::SX; C.sz_ = buf + S::SX + S::SY; C.zs_ = buf + S::SX + S::SY + S::SZ;
        C.col_ = buf + S::SX + S::SY + S::SZ + S::ZS;
        C.ib = lox; C.jb = loy; C.nx = MW; C.np = MW * MH;
        return C;
    }
};
template <typename T, int NZ, int MW, int MH, int NT>
__device__ __forceinline__ CoefS<T> stage_coefs(const LvArgs<T> &L, int lox, int loy, int wcols, int wrows, T *buf) {
    CoefRegs<T, NZ, MW, MH, NT> c;
    c.load(L, lox, loy, wcols, wrows);
    return c.store(buf);
}

// Coefficients of the linearised (eta = 0) level operator at sigma level k: with eta = 0 the sigma
// coefficients (DetermineTransformationConstantsArray.f90:190-201) are (1 - sigma) times per-column
// constants, precomputed at set-up (planes 3, 4, 5 and 0 of col).
template <typename T, class CA>
__device__ __forceinline__ void low_coefs(const CA &C, int i, int j, int k, T &s2, T &s3, T &s4, T &s5) {
    const T om = T(1) - C.zs(k);
    s2 = om * C.col(3, i, j);
    s4 = om * C.col(4, i, j);
    s3 = om * C.col(5, i, j);
    s5 = C.col(0, i, j);
}

// Zebra z-line relaxation of column (i, j): exact tridiagonal solve in sigma of the level operator's
// rows with the neighbouring columns frozen (read through X(k, i, j)); the new column goes to out[0..NZ).
// Couplings to same-colour columns (|di| or |dj| = 2, from the corner-shifted first pass on the sigma
// ghost level) are left out, so that all columns of one colour can be relaxed at once.
// Tridiagonal solve of a zebra column given its rows (lo, di, up, rh; e2 = third entry of the bottom row).
template <typename T, int NZ>
__device__ __forceinline__ void zebra_thomas(T (&lo)[NZ], T (&di)[NZ], T (&up)[NZ], T (&rh)[NZ], T e2, T *out) {
    {   // eliminate the bottom row's third entry with row 1 (no-op when it is zero)
        const T f = e2 != T(0) ? e2 / up[1] : T(0);
        di[0] -= f * lo[1]; up[0] -= f * di[1]; rh[0] -= f * rh[1];
    }
    T inv[NZ];   // reciprocal pivots: one division per level
    inv[0] = T(1) / di[0];
#pragma unroll
    for (int k = 1; k < NZ; k++) {
        const T f = lo[k] * inv[k - 1];
        di[k] -= f * up[k - 1];
        rh[k] -= f * rh[k - 1];
        inv[k] = T(1) / di[k];
    }
    T xn = rh[NZ - 1] * inv[NZ - 1];
    out[NZ - 1] = xn;
#pragma unroll
    for (int k = NZ - 2; k >= 0; k--) { xn = (rh[k] - up[k] * xn) * inv[k]; out[k] = xn; }
}

// zebra_column for a column at least two columns away from every wall (2 <= i <= nxg-3, and the same in
// y): no clamped or shifted stencil windows, so every neighbour term is fixed at compile time.  Same
// terms in the same order as the general case.
template <typename T, int NZ, class XA, class CA>
__device__ __forceinline__ void zebra_column_interior(const LvArgs<T> &L, const CA &C, int i, int j, const XA &X, const T *bc, T *out) {
    const bool d3 = L.dim3;
    const T *SX1 = C.sx(i), *SX2 = SX1 + 3;
    const T *SY1 = C.sy(j), *SY2 = SY1 + 3;
    T lo[NZ], di[NZ], up[NZ], rh[NZ];
    T e2;
    {
        const T *Sz = C.sz(1);
        T s2, s3, s4, s5;
        low_coefs<T>(C, i, j, 1, s2, s3, s4, s5);
        const T hx = C.col(1, i, j), hy = d3 ? C.col(2, i, j) : T(0);
        const T cz = s5 + hx * s2 + hy * s4;
        T a1 = cz * Sz[1], off = T(0);
        off += hx * SX1[0] * X(1, i - 1, j);
        a1 += hx * SX1[1];
        off += hx * SX1[2] * X(1, i + 1, j);
        if (d3) {
            off += hy * SY1[0] * X(1, i, j - 1);
            a1 += hy * SY1[1];
            off += hy * SY1[2] * X(1, i, j + 1);
        }
        lo[0] = T(0); di[0] = cz * Sz[0]; up[0] = a1; e2 = cz * Sz[2];
        rh[0] = bc[0] - off;
    }
#pragma unroll
    for (int k = 1; k < NZ - 1; k++) {
        const T *Z1 = C.sz(k), *Z2 = Z1 + 3;
        T s2, s3, s4, s5;
        low_coefs<T>(C, i, j, k, s2, s3, s4, s5);
        const T czz = s2 * s2 + s4 * s4 + s5 * s5;
        T a[3];
#pragma unroll
        for (int m = 0; m < 3; m++) a[m] = s3 * Z1[m] + czz * Z2[m];
        T off = T(0);
        off += SX2[0] * X(k, i - 1, j);
        a[1] += SX2[1];
        off += SX2[2] * X(k, i + 1, j);
#pragma unroll
        for (int mz = 0; mz < 3; mz++) {
            const int kk = k - 1 + mz;
            const T w0 = T(2) * s2 * Z1[mz] * SX1[0], w1 = T(2) * s2 * Z1[mz] * SX1[1], w2 = T(2) * s2 * Z1[mz] * SX1[2];
            off += w0 * X(kk, i - 1, j);
            a[mz] += w1;
            off += w2 * X(kk, i + 1, j);
        }
        if (d3) {
            off += SY2[0] * X(k, i, j - 1);
            a[1] += SY2[1];
            off += SY2[2] * X(k, i, j + 1);
#pragma unroll
            for (int mz = 0; mz < 3; mz++) {
                const int kk = k - 1 + mz;
                const T w0 = T(2) * s4 * Z1[mz] * SY1[0], w1 = T(2) * s4 * Z1[mz] * SY1[1], w2 = T(2) * s4 * Z1[mz] * SY1[2];
                off += w0 * X(kk, i, j - 1);
                a[mz] += w1;
                off += w2 * X(kk, i, j + 1);
            }
        }
        lo[k] = a[0]; di[k] = a[1]; up[k] = a[2];
        rh[k] = bc[k] - off;
    }
    lo[NZ - 1] = T(0); up[NZ - 1] = T(0); di[NZ - 1] = T(1); rh[NZ - 1] = bc[NZ - 1];
    zebra_thomas<T, NZ>(lo, di, up, rh, e2, out);
}

template <typename T, int NZ, class XA, class CA>
__device__ __forceinline__ void zebra_column(const LvArgs<T> &L, const CA &C, int i, int j, const XA &X, const T *bc, T *out) {
    const int nxg = L.nxg, nyg = L.nyg;
    if (i >= 2 && i <= nxg - 3 && (!L.dim3 || (j >= 2 && j <= nyg - 3))) {
        zebra_column_interior<T, NZ>(L, C, i, j, X, bc, out);
        return;
    }
    const int wsx = cclamp(i - 1, 0, nxg - 3), wsy = L.dim3 ? cclamp(j - 1, 0, nyg - 3) : 0;
    const T *SX1 = C.sx(i), *SX2 = SX1 + 3;
    const T *SY1 = C.sy(j), *SY2 = SY1 + 3;
    int shx = 0; if (wsx == 0) shx = 1; if (wsx + 2 == nxg - 1) shx = -1;
    int shy = 0;
    if (L.dim3) { if (wsy == 0) shy = 1; if (wsy + 2 == nyg - 1) shy = -1; }
    const T *SXc = C.sx(i - shx);
    const T *SYc = C.sy(j - shy);
    T lo[NZ], di[NZ], up[NZ], rh[NZ];
    T e2 = T(0);
#pragma unroll
    for (int k = 0; k < NZ; k++) { lo[k] = T(0); di[k] = T(0); up[k] = T(0); }
    // bottom ghost row
    {
        const T *Sz = C.sz(1);
        T s2, s3, s4, s5;
        low_coefs<T>(C, i, j, 1, s2, s3, s4, s5);
        const T hx = C.col(1, i, j), hy = L.dim3 ? C.col(2, i, j) : T(0);
        const T cz = s5 + hx * s2 + hy * s4;
        T a0 = cz * Sz[0], a1 = cz * Sz[1], a2 = cz * Sz[2], off = T(0);
        // branch-free: the own column's weight goes to the matrix, a neighbour's value times its weight
        // to the right-hand side (the own column is read in place of the neighbour where it is skipped)
#pragma unroll
        for (int m = 0; m < 3; m++) {
            const bool self = wsx + m == i;
            const T v = X(1, self ? i : wsx + m, j);
            a1 += self ? hx * SX1[m] : T(0);
            off += self ? T(0) : hx * SX1[m] * v;
        }
        if (L.dim3) {
#pragma unroll
            for (int m = 0; m < 3; m++) {
                const bool self = wsy + m == j;
                const T v = X(1, i, self ? j : wsy + m);
                a1 += self ? hy * SY1[m] : T(0);
                off += self ? T(0) : hy * SY1[m] * v;
            }
        }
        di[0] = a0; up[0] = a1; e2 = a2;
        rh[0] = bc[0] - off;
    }
#pragma unroll
    for (int k = 1; k < NZ - 1; k++) {
        const T *Z1 = C.sz(k), *Z2 = Z1 + 3;
        T s2, s3, s4, s5;
        low_coefs<T>(C, i, j, k, s2, s3, s4, s5);
        const T czz = s2 * s2 + s4 * s4 + s5 * s5;
        T a[3];
#pragma unroll
        for (int m = 0; m < 3; m++) a[m] = s3 * Z1[m] + czz * Z2[m];
        T off = T(0);
#pragma unroll
        for (int m = 0; m < 3; m++) {
            const bool self = wsx + m == i;
            const T v = X(k, self ? i : wsx + m, j);
            a[1] += self ? SX2[m] : T(0);
            off += self ? T(0) : SX2[m] * v;
        }
#pragma unroll
        for (int mz = 0; mz < 3; mz++) {
            const int kk = k - 1 + mz;
            const int sh = (kk == 0) ? shx : 0;
            const T *S1 = (kk == 0) ? SXc : SX1;
#pragma unroll
            for (int m = 0; m < 3; m++) {
                const T w = T(2) * s2 * Z1[mz] * S1[m];
                const int ii = wsx + sh + m;
                const bool self = ii == i, skip = self || ii - i == 2 || i - ii == 2;
                const T v = X(kk, skip ? i : ii, j);
                a[mz] += self ? w : T(0);
                off += skip ? T(0) : w * v;
            }
        }
        if (L.dim3) {
#pragma unroll
            for (int m = 0; m < 3; m++) {
                const bool self = wsy + m == j;
                const T v = X(k, i, self ? j : wsy + m);
                a[1] += self ? SY2[m] : T(0);
                off += self ? T(0) : SY2[m] * v;
            }
#pragma unroll
            for (int mz = 0; mz < 3; mz++) {
                const int kk = k - 1 + mz;
                const int sh = (kk == 0) ? shy : 0;
                const T *S1 = (kk == 0) ? SYc : SY1;
#pragma unroll
                for (int m = 0; m < 3; m++) {
                    const T w = T(2) * s4 * Z1[mz] * S1[m];
                    const int jj = wsy + sh + m;
                    const bool self = jj == j, skip = self || jj - j == 2 || j - jj == 2;
                    const T v = X(kk, i, skip ? j : jj);
                    a[mz] += self ? w : T(0);
                    off += skip ? T(0) : w * v;
                }
            }
        }
        lo[k] = a[0]; di[k] = a[1]; up[k] = a[2];
        rh[k] = bc[k] - off;
    }
    di[NZ - 1] = T(1); rh[NZ - 1] = bc[NZ - 1];
    zebra_thomas<T, NZ>(lo, di, up, rh, e2, out);
}

// Row k of the zebra line system of column (i, j) (the same rows zebra_column assembles): lo, di, up on
// levels k-1, k, k+1, the right-hand side with the neighbouring columns moved over, and for the bottom
// ghost row (k = 0) its third entry e2 (level 2).  nz = column depth.
template <typename T, class XA, class CA>
__device__ __forceinline__ void zebra_row(const LvArgs<T> &L, const CA &C, int i, int j, int k, int nz, const XA &X, T bk,
                                          T &lo, T &di, T &up, T &rh, T &e2) {
    const int nxg = L.nxg, nyg = L.nyg;
    const bool d3 = L.dim3;
    lo = T(0); up = T(0); e2 = T(0);
    if (k == nz - 1) { di = T(1); rh = bk; return; }
    const int wsx = cclamp(i - 1, 0, nxg - 3), wsy = d3 ? cclamp(j - 1, 0, nyg - 3) : 0;
    const T *SX1 = C.sx(i), *SX2 = SX1 + 3;
    const T *SY1 = C.sy(j), *SY2 = SY1 + 3;
    if (k == 0) {
        const T *Sz = C.sz(1);
        T s2, s3, s4, s5;
        low_coefs<T>(C, i, j, 1, s2, s3, s4, s5);
        const T hx = C.col(1, i, j), hy = d3 ? C.col(2, i, j) : T(0);
        const T cz = s5 + hx * s2 + hy * s4;
        T a1 = cz * Sz[1], off = T(0);
#pragma unroll
        for (int m = 0; m < 3; m++) {
            const bool self = wsx + m == i;
            const T v = X(1, self ? i : wsx + m, j);
            a1 += self ? hx * SX1[m] : T(0);
            off += self ? T(0) : hx * SX1[m] * v;
        }
        if (d3) {
#pragma unroll
            for (int m = 0; m < 3; m++) {
                const bool self = wsy + m == j;
                const T v = X(1, i, self ? j : wsy + m);
                a1 += self ? hy * SY1[m] : T(0);
                off += self ? T(0) : hy * SY1[m] * v;
            }
        }
        di = cz * Sz[0]; up = a1; e2 = cz * Sz[2];
        rh = bk - off;
        return;
    }
    int shx = 0; if (wsx == 0) shx = 1; if (wsx + 2 == nxg - 1) shx = -1;
    int shy = 0;
    if (d3) { if (wsy == 0) shy = 1; if (wsy + 2 == nyg - 1) shy = -1; }
    const T *Z1 = C.sz(k), *Z2 = Z1 + 3;
    T s2, s3, s4, s5;
    low_coefs<T>(C, i, j, k, s2, s3, s4, s5);
    const T czz = s2 * s2 + s4 * s4 + s5 * s5;
    T a[3];
#pragma unroll
    for (int m = 0; m < 3; m++) a[m] = s3 * Z1[m] + czz * Z2[m];
    T off = T(0);
#pragma unroll
    for (int m = 0; m < 3; m++) {
        const bool self = wsx + m == i;
        const T v = X(k, self ? i : wsx + m, j);
        a[1] += self ? SX2[m] : T(0);
        off += self ? T(0) : SX2[m] * v;
    }
#pragma unroll
    for (int mz = 0; mz < 3; mz++) {
        const int kk = k - 1 + mz;
        const int sh = (kk == 0) ? shx : 0;
        const T *S1 = (kk == 0) ? C.sx(i - shx) : SX1;
#pragma unroll
        for (int m = 0; m < 3; m++) {
            const T w = T(2) * s2 * Z1[mz] * S1[m];
            const int ii = wsx + sh + m;
            const bool self = ii == i, skip = self || ii - i == 2 || i - ii == 2;
            const T v = X(kk, skip ? i : ii, j);
            a[mz] += self ? w : T(0);
            off += skip ? T(0) : w * v;
        }
    }
    if (d3) {
#pragma unroll
        for (int m = 0; m < 3; m++) {
            const bool self = wsy + m == j;
            const T v = X(k, i, self ? j : wsy + m);
            a[1] += self ? SY2[m] : T(0);
            off += self ? T(0) : SY2[m] * v;
        }
#pragma unroll
        for (int mz = 0; mz < 3; mz++) {
            const int kk = k - 1 + mz;
            const int sh = (kk == 0) ? shy : 0;
            const T *S1 = (kk == 0) ? C.sy(j - shy) : SY1;
#pragma unroll
            for (int m = 0; m < 3; m++) {
                const T w = T(2) * s4 * Z1[mz] * S1[m];
                const int jj = wsy + sh + m;
                const bool self = jj == j, skip = self || jj - j == 2 || j - jj == 2;
                const T v = X(kk, i, skip ? j : jj);
                a[mz] += self ? w : T(0);
                off += skip ? T(0) : w * v;
            }
        }
    }
    lo = a[0]; di = a[1]; up = a[2];
    rh = bk - off;
}

// zebra_row for a column at least two columns from every wall (no clamped or shifted windows): the
// column's level-0 value is at c in LDS, x- and y-neighbours at fixed offsets SXS and SYS.  Same terms in
// the same order as zebra_row.
template <typename T, int SXS, int SYS, class CA>
__device__ __forceinline__ void zebra_row_interior(const CA &C, bool d3, int i, int j, int k, int nz, const T *c, T bk,
                                                   T &lo, T &di, T &up, T &rh, T &e2) {
    lo = T(0); up = T(0); e2 = T(0);
    if (k == nz - 1) { di = T(1); rh = bk; return; }
    const T *SX1 = C.sx(i), *SX2 = SX1 + 3;
    const T *SY1 = C.sy(j), *SY2 = SY1 + 3;
    const T *xm = c - SXS, *xp = c + SXS, *ym = c - SYS, *yp = c + SYS;
    if (k == 0) {
        const T *Sz = C.sz(1);
        T s2, s3, s4, s5;
        low_coefs<T>(C, i, j, 1, s2, s3, s4, s5);
        const T hx = C.col(1, i, j), hy = d3 ? C.col(2, i, j) : T(0);
        const T cz = s5 + hx * s2 + hy * s4;
        T a1 = cz * Sz[1], off = T(0);
        off += hx * SX1[0] * xm[1];
        a1 += hx * SX1[1];
        off += hx * SX1[2] * xp[1];
        if (d3) {
            off += hy * SY1[0] * ym[1];
            a1 += hy * SY1[1];
            off += hy * SY1[2] * yp[1];
        }
        di = cz * Sz[0]; up = a1; e2 = cz * Sz[2];
        rh = bk - off;
        return;
    }
    const T *Z1 = C.sz(k), *Z2 = Z1 + 3;
    T s2, s3, s4, s5;
    low_coefs<T>(C, i, j, k, s2, s3, s4, s5);
    const T czz = s2 * s2 + s4 * s4 + s5 * s5;
    T a[3];
#pragma unroll
    for (int m = 0; m < 3; m++) a[m] = s3 * Z1[m] + czz * Z2[m];
    T off = T(0);
    off += SX2[0] * xm[k];
    a[1] += SX2[1];
    off += SX2[2] * xp[k];
#pragma unroll
    for (int mz = 0; mz < 3; mz++) {
        const int kk = k - 1 + mz;
        off += T(2) * s2 * Z1[mz] * SX1[0] * xm[kk];
        a[mz] += T(2) * s2 * Z1[mz] * SX1[1];
        off += T(2) * s2 * Z1[mz] * SX1[2] * xp[kk];
    }
    if (d3) {
        off += SY2[0] * ym[k];
        a[1] += SY2[1];
        off += SY2[2] * yp[k];
#pragma unroll
        for (int mz = 0; mz < 3; mz++) {
            const int kk = k - 1 + mz;
            off += T(2) * s4 * Z1[mz] * SY1[0] * ym[kk];
            a[mz] += T(2) * s4 * Z1[mz] * SY1[1];
            off += T(2) * s4 * Z1[mz] * SY1[2] * yp[kk];
        }
    }
    lo = a[0]; di = a[1]; up = a[2];
    rh = bk - off;
}

// Row k (1 <= k <= nz - 2) of an interior column (2 <= i <= nxg - 3, 2 <= j <= nyg - 3) on a uniform
// horizontal grid: zebra_row_interior with the level's constant centred stencils, so each mixed-derivative
// coupling is 2 s2 cx1 sum_m Z1[m] (x[i+1] - x[i-1]) on the three sigma levels around k (the centre
// weight of the first derivative is zero).  c points at the column's level 0 in LDS.
template <typename T, int SXS, int SYS, class CA>
__device__ __forceinline__ void zebra_row_uniform(const LvArgs<T> &L, const CA &C, bool d3, int i, int j, int k, const T *c,
                                                  T bk, T &lo, T &di, T &up, T &rh) {
    const T *Z1 = C.sz(k), *Z2 = Z1 + 3;
    const T om = T(1) - C.zs(k);
    const T s2 = om * C.col(3, i, j), s3 = om * C.col(5, i, j), s5 = C.col(0, i, j);
    const T *xm = c - SXS, *xp = c + SXS;
    const T gx = (Z1[0] * (xp[k - 1] - xm[k - 1]) + Z1[1] * (xp[k] - xm[k])) + Z1[2] * (xp[k + 1] - xm[k + 1]);
    T off = L.cx2 * (xm[k] + xp[k]) + (T(2) * L.cx1) * s2 * gx;
    T czz = s2 * s2 + s5 * s5, dc = L.cxc;
    if (d3) {
        const T s4 = om * C.col(4, i, j);
        const T *ym = c - SYS, *yp = c + SYS;
        const T gy = (Z1[0] * (yp[k - 1] - ym[k - 1]) + Z1[1] * (yp[k] - ym[k])) + Z1[2] * (yp[k + 1] - ym[k + 1]);
        off += L.cy2 * (ym[k] + yp[k]) + (T(2) * L.cy1) * s4 * gy;
        czz += s4 * s4;
        dc += L.cyc;
    }
    lo = s3 * Z1[0] + czz * Z2[0];
    di = (s3 * Z1[1] + czz * Z2[1]) + dc;
    up = s3 * Z1[2] + czz * Z2[2];
    rh = bk - off;
}

// Residual o = b - L x of column (i, j) of a level (all row kinds: interior, bottom ghost, surface,
// wall ghost columns, corners), x read through X(k, i, j).
template <typename T, int NZ, class XA, class CA>
__device__ __forceinline__ void residual_column(const LvArgs<T> &L, const CA &C, int i, int j, const XA &X, const T *bc, T *o) {
    const int nxg = L.nxg, nyg = L.nyg, d3 = L.dim3;
    const bool wx = L.gx && (i == 0 || i == nxg - 1), wy = L.gy && (j == 0 || j == nyg - 1);
    if (wx && wy) {
#pragma unroll
        for (int k = 0; k < NZ; k++) o[k] = bc[k] - X(k, i, j);
    } else if (wx || wy) {
        o[0] = bc[0] - X(0, i, j);
        if (wx) {
            const int ib = i == 0 ? 1 : nxg - 2, ws = cclamp(ib - 1, 0, nxg - 3);
            const T *S = C.sx(ib);
#pragma unroll
            for (int k = 1; k < NZ; k++) o[k] = bc[k] - ((S[0] * X(k, ws, j) + S[1] * X(k, ws + 1, j)) + S[2] * X(k, ws + 2, j));
        } else {
            const int jb = j == 0 ? 1 : nyg - 2, ws = cclamp(jb - 1, 0, nyg - 3);
            const T *S = C.sy(jb);
#pragma unroll
            for (int k = 1; k < NZ; k++) o[k] = bc[k] - ((S[0] * X(k, i, ws) + S[1] * X(k, i, ws + 1)) + S[2] * X(k, i, ws + 2));
        }
    } else {
        const int wsx = cclamp(i - 1, 0, nxg - 3), wsy = d3 ? cclamp(j - 1, 0, nyg - 3) : 0;
        const T *SX1 = C.sx(i), *SX2 = SX1 + 3;
        const T *SY1 = C.sy(j), *SY2 = SY1 + 3;
        T xc[NZ], dxl[NZ], dyl[NZ];
#pragma unroll
        for (int k = 0; k < NZ; k++) {
            xc[k] = X(k, i, j);
            dxl[k] = (SX1[0] * X(k, wsx, j) + SX1[1] * X(k, wsx + 1, j)) + SX1[2] * X(k, wsx + 2, j);
            dyl[k] = d3 ? (SY1[0] * X(k, i, wsy) + SY1[1] * X(k, i, wsy + 1)) + SY1[2] * X(k, i, wsy + 2) : T(0);
        }
        {
            int sh = 0; if (wsx == 0) sh = 1; if (wsx + 2 == nxg - 1) sh = -1;
            if (sh) { const T *S = C.sx(i - sh); dxl[0] = (S[0] * X(0, wsx + sh, j) + S[1] * X(0, wsx + sh + 1, j)) + S[2] * X(0, wsx + sh + 2, j); }
            if (d3) {
                sh = 0; if (wsy == 0) sh = 1; if (wsy + 2 == nyg - 1) sh = -1;
                if (sh) { const T *S = C.sy(j - sh); dyl[0] = (S[0] * X(0, i, wsy + sh) + S[1] * X(0, i, wsy + sh + 1)) + S[2] * X(0, i, wsy + sh + 2); }
            }
        }
        o[NZ - 1] = bc[NZ - 1] - xc[NZ - 1];
#pragma unroll
        for (int k = 1; k < NZ - 1; k++) {
            const T *Z1 = C.sz(k), *Z2 = Z1 + 3;
            const T dz = (Z1[0] * xc[k - 1] + Z1[1] * xc[k]) + Z1[2] * xc[k + 1];
            const T dzz = (Z2[0] * xc[k - 1] + Z2[1] * xc[k]) + Z2[2] * xc[k + 1];
            const T dxx = (SX2[0] * X(k, wsx, j) + SX2[1] * X(k, wsx + 1, j)) + SX2[2] * X(k, wsx + 2, j);
            const T dyy = d3 ? (SY2[0] * X(k, i, wsy) + SY2[1] * X(k, i, wsy + 1)) + SY2[2] * X(k, i, wsy + 2) : T(0);
            const T dxz = (Z1[0] * dxl[k - 1] + Z1[1] * dxl[k]) + Z1[2] * dxl[k + 1];
            const T dyz = (Z1[0] * dyl[k - 1] + Z1[1] * dyl[k]) + Z1[2] * dyl[k + 1];
            T s2, s3, s4, s5;
            low_coefs<T>(C, i, j, k, s2, s3, s4, s5);
            o[k] = bc[k] - (dxx + dyy + s3 * dz + T(2) * (s2 * dxz + s4 * dyz) + (s2 * s2 + s4 * s4 + s5 * s5) * dzz);
        }
        {
            const T *Z1 = C.sz(1);
            const T dz = (Z1[0] * xc[0] + Z1[1] * xc[1]) + Z1[2] * xc[2];
            T s2, s3, s4, s5;
            low_coefs<T>(C, i, j, 1, s2, s3, s4, s5);
            const T hx = C.col(1, i, j), hy = d3 ? C.col(2, i, j) : T(0);
            o[0] = bc[0] - ((s5 + hx * s2 + hy * s4) * dz + hx * dxl[1] + hy * dyl[1]);
        }
    }
}

// Wall ghost point t of a level (t enumerates (ghost column, sigma level) pairs): Neumann rows relaxed
// pointwise, identity rows set to the right-hand side (the role of GaussSeidelSpecialPoints.f90).
template <typename T>
__device__ __forceinline__ void ghost_point(const LvArgs<T> &L, int t, T *__restrict__ x, const T *__restrict__ b) {
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
template <typename T>
__device__ __forceinline__ int ghost_count(const LvArgs<T> &L) {
    return (L.dim3 ? 2 * L.nxg + 2 * (L.nyg - 2) : 2) * L.nzg;
}

// Column index c of the sweep colour -> (i, j); false past the last column of that colour.
template <typename T>
__device__ __forceinline__ bool colour_column(const LvArgs<T> &L, int colour, int c, int &i, int &j) {
    const int nxi = L.nxg - 2 * L.gx, half = (nxi + 1) / 2;
    const int jr = c / half;
    j = L.dim3 ? jr + L.gy : 0;
    if (j >= (L.dim3 ? L.nyg - L.gy : 1)) return false;
    i = L.gx + (((L.gx + j) & 1) != colour ? 1 : 0) + 2 * (c - jr * half);
    return i < L.gx + nxi;
}

// Zebra sweep over the columns of one colour, one thread per column, x read and written in HBM.
template <typename T, int NZ>
__global__ __launch_bounds__(256) void k_mg_zebra(LvArgs<T> L, int colour, T *__restrict__ x, const T *__restrict__ b) {
    if (solve_done(L.skip)) return;
    int i, j;
    if (!colour_column(L, colour, (int)(blockIdx.x * blockDim.x + threadIdx.x), i, j)) return;
    const int nxg = L.nxg;
    auto X = [&](int k, int ii, int jj) -> T { return x[((size_t)jj * nxg + ii) * NZ + k]; };
    const size_t p = (size_t)i + (size_t)nxg * j;
    const CoefG<T> C{L};
    zebra_column<T, NZ>(L, C, i, j, X, b + p * NZ, x + p * NZ);
}

// LDS-tiled zebra z-line relaxation: a block stages a 32 x 8 tile of columns (plus a one-column
// halo) and the tile's right-hand sides, 128 threads each solve the sweep-colour column of one
// (even, odd) pair, and the whole tile is written back coalesced (columns of the other colour
// are rewritten unchanged; nothing else modifies them during the sweep).
template <typename T, int NZ>
__global__ __launch_bounds__(128) void k_mg_zebra_t(LvArgs<T> L, int colour, T *__restrict__ x, const T *__restrict__ b) {
    if (solve_done(L.skip)) return;
    constexpr int TX = 32, TY = 8, WX = TX + 2, WY = TY + 2, NZP = NZ | 1;
    __shared__ T sX[WX * WY * NZP];
    __shared__ T sB[TX * TY * NZP];
    const int nxg = L.nxg, nyg = L.nyg, d3 = L.dim3;
    const int i0 = L.gx + (int)blockIdx.x * TX, j0 = d3 ? L.gy + (int)blockIdx.y * TY : 0;
    const int ie = cmin(i0 + TX, nxg - L.gx), je = d3 ? cmin(j0 + TY, nyg - L.gy) : 1;
    const int lox = i0 - 1, loy = d3 ? j0 - 1 : 0;
    const int wcols = cmin(WX, nxg - lox), wrows = d3 ? cmin(WY, nyg - loy) : 1;
    stage_window<T, NZ, NZP, WY, WX, WX, 128>(sX, x + ((size_t)loy * nxg + lox) * NZ, (size_t)nxg * NZ, wrows, wcols);
    stage_window<T, NZ, NZP, TY, TX, TX, 128>(sB, b + ((size_t)j0 * nxg + i0) * NZ, (size_t)nxg * NZ, je - j0, ie - i0);
    __syncthreads();
    auto X = [&](int k, int ii, int jj) -> T { return sX[((jj - loy) * WX + (ii - lox)) * NZP + k]; };
    const int ty = threadIdx.x / 16, pr = threadIdx.x % 16;
    const int j = j0 + ty;
    int i = i0 + 2 * pr;
    if (((i + j) & 1) != colour) i++;
    if (j < je && i < ie)
        zebra_column<T, NZ>(L, CoefG<T>{L}, i, j, X, sB + ((j - j0) * TX + (i - i0)) * NZP, sX + ((j - loy) * WX + (i - lox)) * NZP);
    __syncthreads();
    for (int r = 0; r < je - j0; r++) {
        T *dst = x + ((size_t)(j0 + r) * nxg + i0) * NZ;
        for (int e = threadIdx.x; e < (ie - i0) * NZ; e += blockDim.x) {
            const int c = e / NZ, k = e - c * NZ;
            dst[e] = sX[((j0 + r - loy) * WX + (i0 + c - lox)) * NZP + k];
        }
    }
}

// One full zebra sweep (red columns, then black columns) in one launch, out of place (xin -> xout: a
// block's staged neighbours must not change while other blocks write their results).  A block owns a
// TX x TY tile of interior columns and stages it with a two-column halo.  Each half-sweep runs in two
// phases: the rows of all columns of the colour are assembled in parallel (one thread per (column,
// level) row, into LDS), then one thread per column runs the tridiagonal solve.  The red half-sweep
// covers the tile and its one-column ring (a red column reads only black neighbours, which do not change
// in the red half-sweep, so the ring gets exactly the values the owning blocks compute); the black
// half-sweep covers the tile and reads those red values.  The wall ghost pass before the sweep is done
// on the staged window and the one after it for the ghost points next to the tile, so one launch does
// (ghost pass, red, black, ghost pass) with x read once and written once.
//
// MODE 1: the first sweep of a preconditioner application, from x = 0 with the right-hand side given as
// the FP64 residual r0: x is not read, b = (T) r0 is staged from r0 and written to bout for the points the
// tile owns (its columns and the ghost points the post-sweep ghost pass gives it), which together cover
// the level; this replaces the separate pass b = (T) r, x = 0 before the V cycle.
// MODE 2: the last sweep of an application: the owned points of the result go to uout as FP64 (the
// preconditioner's output) instead of to xout; this replaces the separate output conversion.
// Prolongation fused into the first post-smoothing sweep of a level with interpolation tables
// (k_mg_prolong_tab's tables and arithmetic; the coarse level is coarsened in x and y, not in sigma).
template <typename T>
struct ProArgs {
    const T *ec;              // coarse correction
    const int2 *pxi, *pyi;    // per fine column / row: left and right coarse index (-1: none)
    const T *pxw, *pyw;       // per fine column / row: left weight, right weight, 1 / sum
    int nxgc;                 // coarse row length
};
// Coarse window of a fine window WX x WY: fine interior column i >= 1 interpolates from coarse columns
// 1 + (i - 1) / 2 and 1 + i / 2.
template <int WX, int WY>
struct ProWindow {
    static constexpr int CWX = WX / 2 + 2, CWY = WY / 2 + 2;
};

template <typename T, int NZ, int TX, int TY, int NT, int MODE = 0>
__global__ __launch_bounds__(NT) void k_mg_zebra_rb(LvArgs<T> L, const T *__restrict__ xin, T *__restrict__ xout, const T *__restrict__ b,
                                                    const double *__restrict__ r0, T *__restrict__ bout, double *__restrict__ uout,
                                                    ProArgs<T> P) {
    // MODE 1: x = 0, b from r0; MODE 2: result to uout in FP64; MODE 3: the coarse correction P is added to
    // x first (prolongation); MODE 4: MODE 3 with the result to uout
    constexpr bool ZX = MODE == 1, ZO = MODE == 2 || MODE == 4, PRO = MODE == 3 || MODE == 4;
    if (!L.spec && solve_done(L.skip)) return;
    constexpr int H = 2, WX = TX + 2 * H, WY = TY + 2 * H, NZP = NZ | 1;
    constexpr int PR = (TX + 2) / 2, NS = PR * (TY + 2);   // red slots: ring-extended tile
    constexpr int CWX = ProWindow<WX, WY>::CWX, CWY = ProWindow<WX, WY>::CWY;
    __shared__ T sX[WX * WY * NZP];
    __shared__ T sC[CoefLds<WX, WY, NZ>::TOTAL];
    __shared__ T sR[4 * NS * NZ];
    __shared__ T sE[NS];
    __shared__ T sB[(TX + 2) * (TY + 2) * NZ];   // right-hand sides of the ring-extended tile
    __shared__ T sP[PRO ? CWX * CWY * NZP : 1];              // coarse correction window
    __shared__ int2 sPi[PRO ? WX + WY : 1];                  // interpolation tables of the window's columns, rows
    __shared__ T sPw[PRO ? 3 * (WX + WY) : 1];
    const int nxg = L.nxg, nyg = L.nyg, d3 = L.dim3;
    // tiles of TX x TY interior columns; when a direction's count leaves a one-column remainder the first
    // tile is one column short instead, so every tile next to a wall has the two columns its ghost
    // values are computed from
    const int nxi = nxg - 2 * L.gx, nyi = d3 ? nyg - 2 * L.gy : 1;
    const int offx = (nxi % TX == 1) ? 1 : 0, offy = (d3 && nyi % TY == 1) ? 1 : 0;
    // (tiles in launch order: the XCD-aware order of k_mg_residual_restrict halves the sweep's HBM fetch at
    // 513^2, 55 -> 29 MB, but the sweep then takes 49.4 instead of 46.0 us)
    const int i0 = L.gx + cmax(0, (int)blockIdx.x * TX - offx), ie = L.gx + cmin(nxi, ((int)blockIdx.x + 1) * TX - offx);
    const int j0 = d3 ? L.gy + cmax(0, (int)blockIdx.y * TY - offy) : 0;
    const int je = d3 ? L.gy + cmin(nyi, ((int)blockIdx.y + 1) * TY - offy) : 1;
    const int lox = cmax(i0 - H, 0), loy = d3 ? cmax(j0 - H, 0) : 0;
    const int wcols = cmin(ie + H, nxg) - lox, wrows = d3 ? cmin(je + H, nyg) - loy : 1;
    // b over the ring-extended tile [i0-1, ie] x [j0-1, je] (clipped to the grid)
    const int blox = i0 - 1, bloy = d3 ? j0 - 1 : 0;
    const int bcols = cmin(ie + 1, nxg) - blox, brows = d3 ? cmin(je + 1, nyg) - bloy : 1;
    // every HBM load of the block (coefficients, x window, b ring) is in flight before the first wait
    CoefRegs<T, NZ, WX, WY, NT> cr;
    WindowRegs<T, NZ, WY, ZX ? 1 : WX, NT> xr;   // unused when ZX
    WindowRegs<typename std::conditional<ZX, double, T>::type, NZ, TY + 2, TX + 2, NT> br;
    WindowRegs<T, NZ, PRO ? CWY : 1, PRO ? CWX : 1, NT> pr;   // unused unless PRO
    cr.load(L, lox, loy, wcols, wrows);
    if constexpr (ZX) br.load(r0 + ((size_t)bloy * nxg + blox) * NZ, (size_t)nxg * NZ, brows, bcols);
    else {
        xr.load(xin + ((size_t)loy * nxg + lox) * NZ, (size_t)nxg * NZ, wrows, wcols);
        br.load(b + ((size_t)bloy * nxg + blox) * NZ, (size_t)nxg * NZ, brows, bcols);
    }
    // coarse columns [cxo, cxo + ccols) x rows [cyo, cyo + crows) reach the window's interior columns
    const int cxo = ((cmax(lox, 1) - 1) >> 1) + 1, cyo = ((cmax(loy, 1) - 1) >> 1) + 1;
    const int ccols = (cmin(lox + wcols - 1, nxg - 2) >> 1) + 2 - cxo, crows = (cmin(loy + wrows - 1, nyg - 2) >> 1) + 2 - cyo;
    int2 pti = make_int2(-1, -1);
    T ptw0 = T(0), ptw1 = T(0), ptw2 = T(0);
    if constexpr (PRO) {
        const int t = (int)threadIdx.x;
        if (t < wcols) {
            pti = P.pxi[lox + t];
            ptw0 = P.pxw[3 * (lox + t)]; ptw1 = P.pxw[3 * (lox + t) + 1]; ptw2 = P.pxw[3 * (lox + t) + 2];
        } else if (t >= WX && t < WX + wrows) {
            pti = P.pyi[loy + t - WX];
            ptw0 = P.pyw[3 * (loy + t - WX)]; ptw1 = P.pyw[3 * (loy + t - WX) + 1]; ptw2 = P.pyw[3 * (loy + t - WX) + 2];
        }
        pr.load(P.ec + ((size_t)cyo * P.nxgc + cxo) * NZ, (size_t)P.nxgc * NZ, crows, ccols);
    }
    if (L.spec && solve_done(L.skip)) return;
    const CoefS<T> C = cr.store(sC);
    if constexpr (ZX) {
        for (int e = threadIdx.x; e < WX * WY * NZP; e += NT) sX[e] = T(0);
    } else {
        xr.template store<NZP, WX>(sX, wrows, wcols);
    }
    br.template store<NZ, TX + 2>(sB, brows, bcols);
    if constexpr (PRO) {
        const int t = (int)threadIdx.x;
        if (t < WX + WY) {
            sPi[t] = pti;
            sPw[3 * t] = ptw0; sPw[3 * t + 1] = ptw1; sPw[3 * t + 2] = ptw2;
        }
        pr.template store<NZP, CWX>(sP, crows, ccols);
    }
    __syncthreads();
    if constexpr (PRO) {
        // x += interpolated coarse correction on the window's interior columns, sigma levels 1 .. NZ - 2
        // (k_mg_prolong_tab's arithmetic), before the ghost pass reads them
        for (int q = threadIdx.x; q < WX * WY; q += NT) {
            const int r = q / WX, c = q - r * WX;
            if (r >= wrows || c >= wcols) continue;
            const int2 xi = sPi[c], yi = sPi[WX + r];
            if (xi.x < 0 || yi.x < 0) continue;
            const T wxl = sPw[3 * c], wxr = sPw[3 * c + 1], sxw = sPw[3 * c + 2];
            const T wyl = sPw[3 * (WX + r)], wyr = sPw[3 * (WX + r) + 1], syw = sPw[3 * (WX + r) + 2];
            const T *ea = sP + ((yi.x - cyo) * CWX + (xi.x - cxo)) * NZP, *eb = sP + ((yi.x - cyo) * CWX + (xi.y - cxo)) * NZP;
            const T *ec = sP + ((yi.y - cyo) * CWX + (xi.x - cxo)) * NZP, *ed = sP + ((yi.y - cyo) * CWX + (xi.y - cxo)) * NZP;
            T *xc = sX + (r * WX + c) * NZP;
#pragma unroll
            for (int k = 1; k < NZ - 1; k++) {
                const T el = (wxl * ea[k] + wxr * eb[k]) * sxw, er = (wxl * ec[k] + wxr * ed[k]) * sxw;
                const T e = (wyl * el + wyr * er) * syw;
                xc[k] = xc[k] + e;
            }
        }
        __syncthreads();
    }
    // owned points [gi0, gi1) x [gj0, gj1): the tile's columns and the ghost points beside them
    const int gi0 = (i0 == L.gx) ? 0 : i0, gi1 = (ie == nxg - L.gx) ? nxg : ie;
    const int gj0 = d3 ? ((j0 == L.gy) ? 0 : j0) : 0, gj1 = d3 ? ((je == nyg - L.gy) ? nyg : je) : 1;
    if constexpr (ZX) {
        const int w = (gi1 - gi0) * NZ;
        for (int r = gj0; r < gj1; r++) {
            T *dst = bout + ((size_t)r * nxg + gi0) * NZ;
            const T *src = sB + ((r - bloy) * (TX + 2) + (gi0 - blox)) * NZ;
            for (int e = threadIdx.x; e < w; e += NT) dst[e] = src[e];
        }
    }
    auto X = [&](int k, int ii, int jj) -> T { return sX[((jj - loy) * WX + (ii - lox)) * NZP + k]; };
    const int jlo = d3 ? L.gy : 0, jhi = d3 ? nyg - L.gy : 1;
    // value of wall ghost point (ii, jj, k) of the ring-extended tile from the staged interior values
    // (GaussSeidelSpecialPoints.f90 rows: identity on the sigma ghost level and at corners, the Neumann row
    // relaxed pointwise elsewhere)
    auto ghost_value = [&](int ii, int jj, int k) -> T {
        const bool wxg = L.gx && (ii == 0 || ii == nxg - 1), wyg = L.gy && (jj == 0 || jj == nyg - 1);
        const T bq = sB[((jj - bloy) * (TX + 2) + (ii - blox)) * NZ + k];
        if ((wxg && wyg) || k == 0) return bq;
        T sum = T(0), dg = T(1);
        if (wxg) {
            const int ib = ii == 0 ? 1 : nxg - 2, ws = cclamp(ib - 1, 0, nxg - 3);
            const T *S = C.sx(ib);
#pragma unroll
            for (int m = 0; m < 3; m++) {
                if (ws + m == ii) dg = S[m];
                else sum += S[m] * X(k, ws + m, jj);
            }
        } else {
            const int jb = jj == 0 ? 1 : nyg - 2, ws = cclamp(jb - 1, 0, nyg - 3);
            const T *S = C.sy(jb);
#pragma unroll
            for (int m = 0; m < 3; m++) {
                if (ws + m == jj) dg = S[m];
                else sum += S[m] * X(k, ii, ws + m);
            }
        }
        return (bq - sum) / dg;
    };
    auto is_ghost = [&](int ii, int jj) -> bool {
        return (L.gx && (ii == 0 || ii == nxg - 1)) || (L.gy && (jj == 0 || jj == nyg - 1));
    };
    // the ghost pass before the sweep, on the ring-extended tile (ghost values depend on interior values
    // only; the sweep reads no ghost point outside that ring: a relaxed ring column's outer neighbour is
    // interior, since no tile is one column or row wide)
    if (blox == 0 || blox + bcols == nxg || (d3 && (bloy == 0 || bloy + brows == nyg))) {
        for (int e = threadIdx.x; e < brows * bcols * NZ; e += NT) {
            const int r = e / (bcols * NZ), rem = e - r * bcols * NZ, c = rem / NZ, k = rem - c * NZ;
            const int ii = blox + c, jj = bloy + r;
            if (is_ghost(ii, jj)) {
                const T v = ghost_value(ii, jj, k);
                sX[((jj - loy) * WX + (ii - lox)) * NZP + k] = v;
            }
        }
        __syncthreads();
    }
    // one half-sweep over the columns of colour COL: red (0) over the ring-extended tile, black (1) over
    // the tile; slot s of row rr is the t-th column of that colour in the row
    auto half = [&](auto colour_tag) {
        constexpr int COL = decltype(colour_tag)::value;
        constexpr int PER = COL == 0 ? PR : TX / 2;
        const int r0 = COL == 0 ? (d3 ? j0 - 1 : 0) : j0;
        const int nr = COL == 0 ? (d3 ? (je - j0) + 2 : 1) : je - j0;
        const int ib = COL == 0 ? i0 - 1 : i0, iend = COL == 0 ? ie : ie - 1;
        const int ns = PER * nr;
        auto slot = [&](int sl, int &i, int &j) -> bool {
            const int rr = sl / PER, t = sl - rr * PER;
            j = r0 + rr;
            i = ib + (((ib + j) & 1) != COL ? 1 : 0) + 2 * t;
            return i <= iend && j >= jlo && j < jhi && i >= L.gx && i < nxg - L.gx;
        };
        // rows in level-major order (task = k * NSM + slot): the lanes of a wave share the level, so
        // bottom, interior and surface rows do not diverge inside a wave; row arrays are slot-fastest
        // (sR[(m * NZ + k) * NS + slot]), conflict-free for both phases
        constexpr int NSM = PER * (COL == 0 ? TY + 2 : TY);
        for (int task = threadIdx.x; task < NSM * NZ; task += NT) {
            const int k = task / NSM, sl = task - k * NSM;
            int i, j;
            if (sl >= ns || !slot(sl, i, j)) continue;
            T lo, di, up, rh, e2;
            const T bk = sB[((j - bloy) * (TX + 2) + (i - blox)) * NZ + k];
            if (i >= 2 && i <= nxg - 3 && (!d3 || (j >= 2 && j <= nyg - 3))) {
                const T *cc = sX + ((j - loy) * WX + (i - lox)) * NZP;
                if (L.uni && k >= 1 && k <= NZ - 2) {
                    zebra_row_uniform<T, NZP, WX * NZP>(L, C, d3, i, j, k, cc, bk, lo, di, up, rh);
                    e2 = T(0);
                } else {
                    zebra_row_interior<T, NZP, WX * NZP>(C, d3, i, j, k, NZ, cc, bk, lo, di, up, rh, e2);
                }
            } else
                zebra_row<T>(L, C, i, j, k, NZ, X, bk, lo, di, up, rh, e2);
            sR[(0 * NZ + k) * NS + sl] = lo; sR[(1 * NZ + k) * NS + sl] = di;
            sR[(2 * NZ + k) * NS + sl] = up; sR[(3 * NZ + k) * NS + sl] = rh;
            if (k == 0) sE[sl] = e2;
        }
        __syncthreads();
        for (int sl = threadIdx.x; sl < ns; sl += NT) {
            int i, j;
            if (!slot(sl, i, j)) continue;
            T lo[NZ], di[NZ], up[NZ], rh[NZ];
#pragma unroll
            for (int k = 0; k < NZ; k++) {
                lo[k] = sR[(0 * NZ + k) * NS + sl]; di[k] = sR[(1 * NZ + k) * NS + sl];
                up[k] = sR[(2 * NZ + k) * NS + sl]; rh[k] = sR[(3 * NZ + k) * NS + sl];
            }
            zebra_thomas<T, NZ>(lo, di, up, rh, sE[sl], sX + ((j - loy) * WX + (i - lox)) * NZP);
        }
        __syncthreads();
    };
    half(std::integral_constant<int, 0>{});
    half(std::integral_constant<int, 1>{});
    using O = typename std::conditional<ZO, double, T>::type;
    O *const out = ZO ? (O *)uout : (O *)xout;
    for (int r = 0; r < je - j0; r++)
        store_row<NZ, NZP, NT>(out + ((size_t)(j0 + r) * nxg + i0) * NZ, sX + ((j0 + r - loy) * WX + (i0 - lox)) * NZP, ie - i0);
    // the ghost pass after the sweep for the ghost points this tile owns: wall columns beside its rows,
    // wall rows beside its columns, and the corners of the tiles at the corners of the domain
    {
        if (gi0 != i0 || gi1 != ie || gj0 != j0 || gj1 != je) {
            const int w = gi1 - gi0;
            for (int e = threadIdx.x; e < (gj1 - gj0) * w * NZ; e += NT) {
                const int r = e / (w * NZ), rem = e - r * w * NZ, c = rem / NZ, k = rem - c * NZ;
                const int ii = gi0 + c, jj = gj0 + r;
                if (is_ghost(ii, jj)) out[((size_t)jj * nxg + ii) * NZ + k] = (O)ghost_value(ii, jj, k);
            }
        }
    }
}

template <typename T>
__global__ void k_mg_ghost_p(LvArgs<T> L, T *__restrict__ x, const T *__restrict__ b) {
    if (solve_done(L.skip)) return;
    ghost_point(L, (int)(blockIdx.x * blockDim.x + threadIdx.x), x, b);
}

// LDS-tiled residual r = b - L x (16 x TY columns per block, halo 2 on domain ends as the matvec; x-z
// problems: one row of 16 * TY columns), level coefficients staged in LDS with the tile.
template <typename T, int NZ, int TY>
__global__ __launch_bounds__(16 * TY) void k_mg_residual_t(LvArgs<T> L, const T *__restrict__ x, const T *__restrict__ b, T *__restrict__ r) {
    if (!L.spec && solve_done(L.skip)) return;
    constexpr int TX = 16, R = 2, WX = TX + 2 * R, WY = TY + 2 * R, NZP = NZ | 1, NT = TX * TY;
    constexpr int RS2 = TX * TY + 2 * R;       // x-z problems: the tile is one row of TX * TY columns
    constexpr int LX = WX * WY > RS2 ? WX * WY : RS2;
    __shared__ T sX[LX * NZP];
    __shared__ T sC[CoefLds<WX, WY, NZ>::TOTAL > CoefLds<RS2, 1, NZ>::TOTAL ? CoefLds<WX, WY, NZ>::TOTAL : CoefLds<RS2, 1, NZ>::TOTAL];
    const int nxg = L.nxg, nyg = L.nyg, d3 = L.dim3;
    const int i0 = d3 ? (int)blockIdx.x * TX : (int)blockIdx.x * (TX * TY), j0 = d3 ? (int)blockIdx.y * TY : 0;
    const int tw = d3 ? TX : TX * TY;        // columns per tile row
    const int lox = cmax(cmin(i0 - R, nxg - 4), 0);
    const int loy = d3 ? cmax(cmin(j0 - R, nyg - 4), 0) : 0;
    const int wcols = cmin(tw + 2 * R, nxg - lox), wrows = d3 ? cmin(WY, nyg - loy) : 1;
    const int rowstride = d3 ? WX : RS2;
    const int tx = d3 ? (int)threadIdx.x % TX : (int)threadIdx.x, ty = d3 ? (int)threadIdx.x / TX : 0;
    const int i = i0 + tx, j = j0 + ty;
    const bool active = i < nxg && j < nyg;
    // every HBM load of the block (coefficients, x window, the thread's column of b) is in flight before
    // the first wait
    T bc[NZ];
    const T *bp = b + ((size_t)i + (size_t)nxg * j) * NZ;
#pragma unroll
    for (int k = 0; k < NZ; k++) bc[k] = active ? bp[k] : T(0);
    CoefS<T> C;
    if (d3) {
        CoefRegs<T, NZ, WX, WY, NT> cr;
        WindowRegs<T, NZ, WY, WX, NT> xr;
        cr.load(L, lox, loy, wcols, wrows);
        xr.load(x + ((size_t)loy * nxg + lox) * NZ, (size_t)nxg * NZ, wrows, wcols);
        if (L.spec && solve_done(L.skip)) return;
        C = cr.store(sC);
        xr.template store<NZP, WX>(sX, wrows, wcols);
    } else {
        CoefRegs<T, NZ, RS2, 1, NT> cr;
        WindowRegs<T, NZ, 1, RS2, NT> xr;
        cr.load(L, lox, loy, wcols, wrows);
        xr.load(x + (size_t)lox * NZ, 0, 1, wcols);
        if (L.spec && solve_done(L.skip)) return;
        C = cr.store(sC);
        xr.template store<NZP, RS2>(sX, 1, wcols);
    }
    __syncthreads();
    auto X = [&](int k, int ii, int jj) -> T { return sX[((jj - loy) * rowstride + (ii - lox)) * NZP + k]; };
    T o[NZ];
    if (active) residual_column<T, NZ>(L, C, i, j, X, bc, o);
    __syncthreads();
    if (active) {
        T *dst = sX + ((j - loy) * rowstride + (i - lox)) * NZP;
#pragma unroll
        for (int k = 0; k < NZ; k++) dst[k] = o[k];
    }
    __syncthreads();
    const int ie = cmin(i0 + tw, nxg), je = d3 ? cmin(j0 + TY, nyg) : 1;
    for (int rr = 0; rr < je - j0; rr++)
        store_row<NZ, NZP, NT>(r + ((size_t)(j0 + rr) * nxg + i0) * NZ, sX + ((j0 + rr - loy) * rowstride + (i0 - lox)) * NZP, ie - i0);
}

// Restriction (MGRestriction3DExplicit.f90): separable full weighting x -> y -> z on interior
// points below the free surface, then injection; coarse ghost entries are zero.
struct RestrictGeo {
    int nxgf, nygf, nzgf, nxf, nyf, nzf, gx, gy;     // fine
    int nxgc, nygc, nzgc, nxc, nyc, nzc;             // coarse
    int xc, yc, zc;
    const double *xf, *yf, *zf;                      // fine coordinates
    const int *skip;                                 // solve_done flag (nullptr: always run)
    int spec;                                        // 1: loads are issued before skip is tested
};
// Restricted residual at coarse point (kc, ic, jc) (0 on ghost points).
template <typename T>
__device__ __forceinline__ T restrict_at(const RestrictGeo &G, int kc, int ic, int jc, const T *__restrict__ rf) {
    const int icn = ic - G.gx, jcn = jc - G.gy;
    if (kc < 1 || icn < 0 || icn >= G.nxc || jcn < 0 || jcn >= G.nyc) return T(0);
    const int kf = 1 + (kc - 1) * (1 + G.zc), ifn = G.gx + icn * (1 + G.xc), jfn = G.gy + jcn * (1 + G.yc);
    auto R = [&](int k, int i, int j) -> T { return rf[(size_t)(j * G.nxgf + i) * G.nzgf + k]; };
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
            i >= G.gx && i <= G.nxf + G.gx - 1 && (((i - G.gx) & G.xc) == 0)) {
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
    return v;
}
// Restricted residual at coarse point q < nxgc * nygc * nzgc.
template <typename T>
__device__ __forceinline__ T restrict_value(const RestrictGeo &G, size_t q, const T *__restrict__ rf) {
    const int kc = (int)(q % G.nzgc);
    const size_t col = q / G.nzgc;
    return restrict_at(G, kc, (int)(col % G.nxgc), (int)(col / G.nxgc), rf);
}
// Coarse point q; also clears the coarse iterate xz (the coarse V cycle starts from zero).
template <typename T>
__device__ __forceinline__ void restrict_point(const RestrictGeo &G, size_t q, const T *__restrict__ rf, T *__restrict__ rc, T *__restrict__ xz) {
    const size_t nc = (size_t)G.nxgc * G.nygc * G.nzgc;
    if (q >= nc) return;
    xz[q] = T(0);
    rc[q] = restrict_value(G, q, rf);
}
template <typename T>
__global__ void k_mg_restrict(RestrictGeo G, const T *__restrict__ rf, T *__restrict__ rc, T *__restrict__ xz) {
    if (!G.spec && solve_done(G.skip)) return;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)G.nxgc * G.nygc * G.nzgc) return;
    const T v = restrict_value(G, q, rf);
    if (G.spec && solve_done(G.skip)) return;
    xz[q] = T(0);
    rc[q] = v;
}
// The same with one grid row of coarse columns per blockIdx.y and the coarse level count NZC known at
// compile time: the point decodes with no runtime division.
template <typename T, int NZC>
__global__ void k_mg_restrict_t(RestrictGeo G, const T *__restrict__ rf, T *__restrict__ rc, T *__restrict__ xz) {
    if (!G.spec && solve_done(G.skip)) return;
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= G.nxgc * NZC) return;
    const int ic = t / NZC, kc = t - ic * NZC, jc = (int)blockIdx.y;
    const T v = restrict_at(G, kc, ic, jc, rf);
    if (G.spec && solve_done(G.skip)) return;
    const size_t q = (size_t)jc * G.nxgc * NZC + t;
    xz[q] = T(0);
    rc[q] = v;
}

// Prolongation (MGProlongation3DExplicit.f90) of the coarse correction, added to x on interior
// columns below the free surface.
// Interpolated correction at fine point q < nxgf * nygf * nzgf; false where q takes none (ghost points,
// the sigma ghost level and the surface).
// (The factors XF, YF, ZF are 1 or 2: x % F is x & (F - 1) and x / F is x >> (F - 1) on the
// non-negative offsets they are applied to.)
template <typename T>
__device__ __forceinline__ bool prolong_at(const RestrictGeo &G, int k, int i, int j, const T *__restrict__ ec, T &out) {
    if (k < 1 || k > G.nzf - 1 || i < G.gx || i > G.nxf + G.gx - 1 || j < G.gy || j > G.nyf + G.gy - 1) return false;
    const int XF = G.xc ? 2 : 1, YF = G.yc ? 2 : 1, ZF = G.zc ? 2 : 1;
    const int xs = G.xc ? 1 : 0, ys = G.yc ? 1 : 0, zs = G.zc ? 1 : 0;
    auto E0 = [&](int kk, int ii, int jj) -> T {
        if (kk < 1 || kk > G.nzf - 1 || ((kk - 1) & zs) != 0) return T(0);
        if (((ii - G.gx) & xs) != 0 || ((jj - G.gy) & ys) != 0) return T(0);
        const int kc = 1 + ((kk - 1) >> zs), ic = G.gx + ((ii - G.gx) >> xs), jc = G.gy + ((jj - G.gy) >> ys);
        if (kc > G.nzc - 1) return T(0);
        return ec[(size_t)(jc * G.nxgc + ic) * G.nzgc + kc];
    };
    auto EX = [&](int kk, int ii, int jj) -> T {
        if (XF == 2 && ((jj - G.gy) & ys) == 0 && ((ii - G.gx) % 2) == 1 && kk >= 1 && kk <= G.nzf - 1 && ((kk - 1) & zs) == 0) {
            const T dW = (T)(G.xf[ii] - G.xf[ii - 1]), dE = (T)(G.xf[ii + 1] - G.xf[ii]);
            const T dxf = T(1) / (dW + dE);
            return (dE * E0(kk, ii - 1, jj) + dW * E0(kk, ii + 1, jj)) * dxf;
        }
        return E0(kk, ii, jj);
    };
    auto EY = [&](int kk, int ii, int jj) -> T {
        if (kk > G.nzf - 1) return T(0);
        if (YF == 2 && ((jj - G.gy) % 2) == 1 && kk >= 1 && ((kk - 1) & zs) == 0) {
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
    out = e;
    return true;
}
template <typename T>
__device__ __forceinline__ bool prolong_value(const RestrictGeo &G, size_t q, const T *__restrict__ ec, T &out) {
    const int k = (int)(q % G.nzgf);
    const size_t col = q / G.nzgf;
    return prolong_at(G, k, (int)(col % G.nxgf), (int)(col / G.nxgf), ec, out);
}
template <typename T>
__device__ __forceinline__ void prolong_point(const RestrictGeo &G, size_t q, const T *__restrict__ ec, T *__restrict__ xf) {
    if (q >= (size_t)G.nxgf * G.nygf * G.nzgf) return;
    T e;
    if (prolong_value(G, q, ec, e)) xf[q] = xf[q] + e;
}
template <typename T>
__global__ void k_mg_prolong_add(RestrictGeo G, const T *__restrict__ ec, T *__restrict__ xf) {
    if (!G.spec && solve_done(G.skip)) return;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)G.nxgf * G.nygf * G.nzgf) return;
    T e;
    if (!prolong_value(G, q, ec, e)) return;
    const T x0 = xf[q];
    if (G.spec && solve_done(G.skip)) return;
    xf[q] = x0 + e;
}
// The same with one grid row of fine columns per blockIdx.y and the fine level count NZF known at
// compile time.
template <typename T, int NZF>
__global__ void k_mg_prolong_add_t(RestrictGeo G, const T *__restrict__ ec, T *__restrict__ xf) {
    if (!G.spec && solve_done(G.skip)) return;
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= G.nxgf * NZF) return;
    const int i = t / NZF, k = t - i * NZF, j = (int)blockIdx.y;
    T e;
    if (!prolong_at(G, k, i, j, ec, e)) return;
    const size_t q = (size_t)j * G.nxgf * NZF + t;
    const T x0 = xf[q];
    if (G.spec && solve_done(G.skip)) return;
    xf[q] = x0 + e;
}

// Restriction onto a coarse level coarsened in x and y but not in sigma (kc = kf), with the 3-point weights
// of each coarse column and row from tables (restrict_at's values: the fine residual smoothed along x on
// the coarse point's row and its two neighbour rows, injected on wall ghost rows, then along y; injected on
// the top level; zero on coarse ghost points).  Also clears the coarse iterate.
template <typename T, int NZC>
__global__ void k_mg_restrict_tab(RestrictGeo G, const T *__restrict__ rxw, const T *__restrict__ ryw, const T *__restrict__ rf,
                                  T *__restrict__ rc, T *__restrict__ xz) {
    if (!G.spec && solve_done(G.skip)) return;
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= G.nxgc * NZC) return;
    const int ic = t / NZC, k = t - ic * NZC, jc = (int)blockIdx.y;
    const int icn = ic - G.gx, jcn = jc - G.gy;
    T v = T(0);
    if (k >= 1 && icn >= 0 && icn < G.nxc && jcn >= 0 && jcn < G.nyc) {
        const int i = G.gx + 2 * icn, j = G.gy + 2 * jcn;
        auto R = [&](int ii, int jj) -> T { return rf[((size_t)jj * G.nxgf + ii) * NZC + k]; };
        if (k <= G.nzf - 1) {
            const T x0 = rxw[4 * ic], x1 = rxw[4 * ic + 1], x2 = rxw[4 * ic + 2], ax = rxw[4 * ic + 3];
            auto Xv = [&](int jj) -> T {
                if (jj < G.gy || jj > G.nyf + G.gy - 1) return R(i, jj);
                return ax * ((x0 * R(i - 1, jj) + x1 * R(i, jj)) + x2 * R(i + 1, jj));
            };
            const T y0 = ryw[4 * jc], y1 = ryw[4 * jc + 1], y2 = ryw[4 * jc + 2], ay = ryw[4 * jc + 3];
            v = ay * ((y0 * Xv(j - 1) + y1 * Xv(j)) + y2 * Xv(j + 1));
        } else {
            v = R(i, j);
        }
    }
    if (G.spec && solve_done(G.skip)) return;
    const size_t q = (size_t)jc * G.nxgc * NZC + t;
    xz[q] = T(0);
    rc[q] = v;
}

// LDS elements of k_mg_residual_restrict's window (x, then the residual tile over it), coefficients and
// restriction weights.
template <int NZ, int CX, int CY>
struct ResRestrictLds {
    static constexpr int FW = 2 * CX + 1, FH = 2 * CY + 1, WX = FW + 4, WY = FH + 4, NZP = NZ | 1;
    static constexpr int X = WX * WY * NZP, C = CoefLds<WX, WY, NZ>::TOTAL, W = 4 * (CX + CY);
    static constexpr int TOTAL = X + C + W;
};
// Residual of a level and its restriction onto a coarse level coarsened in x and y but not in sigma, in one
// launch (k_mg_residual_t followed by k_mg_restrict_tab, same arithmetic).  A block owns CX x CY coarse
// columns (ghost columns included); it computes the fine residual (residual_column) on the (2 CX + 1) x
// (2 CY + 1) fine columns their restriction stencils reach, one thread per column (the fine column between
// two tiles is computed by both), keeps it in LDS over the staged x window and restricts from there, so the
// fine residual never goes to HBM.  Also clears the coarse iterate.  Requires wall ghost layers in x and y
// (gx = gy = 1), as every level with restriction tables has.
template <typename T, int NZ, int CX, int CY, int NT>
__global__ __launch_bounds__(NT) void k_mg_residual_restrict(LvArgs<T> L, RestrictGeo G, const T *__restrict__ rxw, const T *__restrict__ ryw,
                                                             const T *__restrict__ x, const T *__restrict__ b, T *__restrict__ rc,
                                                             T *__restrict__ xz) {
    if (!L.spec && solve_done(L.skip)) return;
    using S = ResRestrictLds<NZ, CX, CY>;
    constexpr int FW = S::FW, FH = S::FH, WX = S::WX, WY = S::WY, NZP = S::NZP, R = 2;
    static_assert(FW * FH <= NT, "one thread per fine column of the tile");
    static_assert(4 * (CX + CY) <= NT, "one weight per thread");
    __shared__ T sX[S::X];
    __shared__ T sC[S::C];
    __shared__ T sW[S::W];
    const int nxg = L.nxg, nyg = L.nyg, t = (int)threadIdx.x;
    // XCD-aware tile order (xcd_tile over the row-major tile index): each XCD takes one contiguous band of
    // tile rows, so the halos neighbouring tiles share are re-read from that XCD's L2 (at 513^2: HBM fetch
    // 57 -> 29 MB, 22.2 -> 20.0 us)
    const int tb = xcd_tile((int)(blockIdx.y * gridDim.x + blockIdx.x), (int)(gridDim.x * gridDim.y));
    const int ic0 = (tb % (int)gridDim.x) * CX, jc0 = (tb / (int)gridDim.x) * CY;
    // fine column of coarse column ic is 2 ic - 1; the tile's stencils reach fine columns fi0 .. fi0 + FW - 1
    const int fi0 = 2 * ic0 - 2, fj0 = 2 * jc0 - 2;
    const int lox = cmax(cmin(fi0 - R, nxg - 4), 0), loy = cmax(cmin(fj0 - R, nyg - 4), 0);
    const int wcols = cmin(WX, nxg - lox), wrows = cmin(WY, nyg - loy);
    const int tx = t % FW, ty = t / FW;
    const int i = fi0 + tx, j = fj0 + ty;
    const bool active = t < FW * FH && i >= 0 && i < nxg && j >= 0 && j < nyg;
    // every HBM load of the block (coefficients, x window, the thread's column of b, the tile's restriction
    // weights) is in flight before the first wait
    T bc[NZ];
    const T *bp = b + (active ? (size_t)i + (size_t)nxg * j : (size_t)0) * NZ;
#pragma unroll
    for (int k = 0; k < NZ; k++) bc[k] = active ? bp[k] : T(0);
    T w = T(0);
    if (t < 4 * CX) {
        if (ic0 + t / 4 < G.nxgc) w = rxw[4 * ic0 + t];
    } else if (t < 4 * (CX + CY)) {
        if (jc0 + (t - 4 * CX) / 4 < G.nygc) w = ryw[4 * jc0 + t - 4 * CX];
    }
    CoefRegs<T, NZ, WX, WY, NT> cr;
    WindowRegs<T, NZ, WY, WX, NT> xr;
    cr.load(L, lox, loy, wcols, wrows);
    xr.load(x + ((size_t)loy * nxg + lox) * NZ, (size_t)nxg * NZ, wrows, wcols);
    if (L.spec && solve_done(L.skip)) return;
    const CoefS<T> C = cr.store(sC);
    xr.template store<NZP, WX>(sX, wrows, wcols);
    if (t < 4 * (CX + CY)) sW[t] = w;
    __syncthreads();
    auto X = [&](int k, int ii, int jj) -> T { return sX[((jj - loy) * WX + (ii - lox)) * NZP + k]; };
    T o[NZ];
    if (active) residual_column<T, NZ>(L, C, i, j, X, bc, o);
    __syncthreads();
    if (active) {
        T *dst = sX + (ty * FW + tx) * NZP;
#pragma unroll
        for (int k = 0; k < NZ; k++) dst[k] = o[k];
    }
    __syncthreads();
    // restriction (k_mg_restrict_tab's arithmetic) of the tile's coarse points, CX * NZ contiguous values
    // per coarse row
    for (int e = t; e < CY * CX * NZ; e += NT) {
        const int rr = e / (CX * NZ), c = e - rr * (CX * NZ), cc = c / NZ, k = c - cc * NZ;
        const int ic = ic0 + cc, jc = jc0 + rr;
        if (ic >= G.nxgc || jc >= G.nygc) continue;
        const int icn = ic - 1, jcn = jc - 1;
        T v = T(0);
        if (k >= 1 && icn >= 0 && icn < G.nxc && jcn >= 0 && jcn < G.nyc) {
            const int ti = 2 * cc + 1, tj = 2 * rr + 1;   // tile position of fine column (2 ic - 1, 2 jc - 1)
            auto Rt = [&](int a, int bb) -> T { return sX[(bb * FW + a) * NZP + k]; };
            if (k <= G.nzf - 1) {
                const T x0 = sW[4 * cc], x1 = sW[4 * cc + 1], x2 = sW[4 * cc + 2], ax = sW[4 * cc + 3];
                const int jf = 2 * jc - 1;
                auto Xv = [&](int dj) -> T {
                    const int jj = jf + dj;
                    if (jj < 1 || jj > G.nyf) return Rt(ti, tj + dj);
                    return ax * ((x0 * Rt(ti - 1, tj + dj) + x1 * Rt(ti, tj + dj)) + x2 * Rt(ti + 1, tj + dj));
                };
                const T *Y = sW + 4 * CX + 4 * rr;
                v = Y[3] * ((Y[0] * Xv(-1) + Y[1] * Xv(0)) + Y[2] * Xv(1));
            } else {
                v = Rt(ti, tj);
            }
        }
        const size_t q = ((size_t)jc * G.nxgc + ic) * NZ + k;
        xz[q] = T(0);
        rc[q] = v;
    }
}

// The same prolongation for a coarse level coarsened in x and y but not in sigma (kc = k), from per-column
// and per-row tables (left and right coarse index, weights left, right and 1 / sum, as prolong_at forms
// them): e = (wyl (wxl a + wxr b) sx + wyr (wxl c + wxr d) sx) sy over the (up to) four coarse neighbours,
// which is prolong_at's value with weights (1, 0, 1) on the coarse lines.  No branch on the point's parity.
template <typename T, int NZF>
__global__ void k_mg_prolong_tab(RestrictGeo G, const int2 *__restrict__ pxi, const T *__restrict__ pxw, const int2 *__restrict__ pyi,
                                 const T *__restrict__ pyw, const T *__restrict__ ec, T *__restrict__ xf) {
    if (!G.spec && solve_done(G.skip)) return;
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= G.nxgf * NZF) return;
    const int i = t / NZF, k = t - i * NZF, j = (int)blockIdx.y;
    const int2 yi = pyi[j], xi = pxi[i];
    if (yi.x < 0 || xi.x < 0 || k < 1 || k > G.nzf - 1) return;
    const T wxl = pxw[3 * i], wxr = pxw[3 * i + 1], sx = pxw[3 * i + 2];
    const T wyl = pyw[3 * j], wyr = pyw[3 * j + 1], sy = pyw[3 * j + 2];
    const size_t rl = (size_t)yi.x * G.nxgc, rr = (size_t)yi.y * G.nxgc;
    const T a = ec[(rl + xi.x) * NZF + k], b = ec[(rl + xi.y) * NZF + k];
    const T c = ec[(rr + xi.x) * NZF + k], d = ec[(rr + xi.y) * NZF + k];
    const T el = (wxl * a + wxr * b) * sx, er = (wxl * c + wxr * d) * sx;
    const T e = (wyl * el + wyr * er) * sy;
    const size_t q = (size_t)j * G.nxgf * NZF + t;
    const T x0 = xf[q];
    if (G.spec && solve_done(G.skip)) return;
    xf[q] = x0 + e;
}

// Coarsest level: x = Minv b with the dense inverse computed at set-up, stored as M (T, or FP16 with
// sums in T) in rows of `pitch` elements (a multiple of 8; zeros past n); b holds pitch elements (zeros
// past n).  One 64-lane wave per row: each lane reads 16 bytes of the row and the matching b values per
// step (vector loads), sums those products in order, and the lanes' sums are reduced in a fixed order.
template <typename T, typename M>
__global__ __launch_bounds__(256) void k_mg_coarse(int n, int pitch, const M *__restrict__ Minv, const T *__restrict__ b,
                                                   T *__restrict__ x, const int *skip) {
    if (solve_done(skip)) return;
    constexpr int V = 16 / (int)sizeof(M);
    const int lane = threadIdx.x & 63, r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= n) return;
    const M *row = Minv + (size_t)r * pitch;
    T s = T(0);
#pragma unroll 4
    for (int c0 = lane * V; c0 < n; c0 += 64 * V) {
        const uint4 q = *reinterpret_cast<const uint4 *>(row + c0);
        M m[V];
        __builtin_memcpy(m, &q, 16);
        T bv[V];
        __builtin_memcpy(bv, __builtin_assume_aligned(b + c0, 16), V * sizeof(T));
        T p = T(0);
#pragma unroll
        for (int v = 0; v < V; v++) p += (T)m[v] * bv[v];
        s += p;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) x[r] = s;
}


// The bottom of the V cycle in one workgroup: levels 0 .. nlev-1 of a hierarchy (0 = coarsest) whose
// finest member `top` = nlev-1 arrives with its right-hand side set and x = 0.  Runs the same sequence
// as the multi-kernel V cycle (ghost pass, nu1 x (red, black, ghost), residual, restriction, ..., dense
// coarse solve, ..., prolongation, ghost pass, nu2 x (red, black, ghost)), with the level vectors in HBM /
// L2 and a workgroup barrier between the steps.  Small levels cost one barrier per step instead of one
// kernel launch.
constexpr int MG_BOTTOM_MAX = 10;
template <typename T>
struct BottomArgs {
    int nlev, nu1, nu2, n0, pitch0;   // pitch0: row pitch (elements) of minv
    const T *minv;
    LvArgs<T> L[MG_BOTTOM_MAX];
    T *x[MG_BOTTOM_MAX];
    T *b[MG_BOTTOM_MAX];
    T *r[MG_BOTTOM_MAX];
    RestrictGeo G[MG_BOTTOM_MAX];   // G[l]: level l (fine) -> level l-1 (coarse), l >= 1
    const int *skip;                // solve_done flag (nullptr: always run)
};

template <typename T, int NZ>
__device__ void bottom_sweep(const LvArgs<T> &L, int colour, T *__restrict__ x, const T *__restrict__ b) {
    const int nxi = L.nxg - 2 * L.gx, half = (nxi + 1) / 2, nyi = L.dim3 ? L.nyg - 2 * L.gy : 1;
    const int nxg = L.nxg;
    auto X = [&](int k, int ii, int jj) -> T { return x[((size_t)jj * nxg + ii) * NZ + k]; };
    for (int c = threadIdx.x; c < half * nyi; c += blockDim.x) {
        int i, j;
        if (!colour_column(L, colour, c, i, j)) continue;
        const size_t p = (size_t)i + (size_t)nxg * j;
        zebra_column<T, NZ>(L, CoefG<T>{L}, i, j, X, b + p * NZ, x + p * NZ);
    }
}

template <typename T, int NZ>
__device__ void bottom_residual(const LvArgs<T> &L, const T *__restrict__ x, const T *__restrict__ b, T *__restrict__ r) {
    const int nxg = L.nxg;
    auto X = [&](int k, int ii, int jj) -> T { return x[((size_t)jj * nxg + ii) * NZ + k]; };
    for (int c = threadIdx.x; c < L.nxy; c += blockDim.x) {
        const int i = c % nxg, j = c / nxg;
        T o[NZ];
        residual_column<T, NZ>(L, CoefG<T>{L}, i, j, X, b + (size_t)c * NZ, o);
#pragma unroll
        for (int k = 0; k < NZ; k++) r[(size_t)c * NZ + k] = o[k];
    }
}

template <typename T>
__device__ void bottom_ghost(const LvArgs<T> &L, T *x, const T *b) {
    const int nb = ghost_count(L);
    for (int t = threadIdx.x; t < nb; t += blockDim.x) ghost_point(L, t, x, b);
}

template <typename T, int NZ>
__device__ void bottom_smooth(const LvArgs<T> &L, int nu, T *x, const T *b) {
    bottom_ghost(L, x, b);
    __syncthreads();
    for (int it = 0; it < nu; it++) {
        bottom_sweep<T, NZ>(L, 0, x, b);
        __syncthreads();
        bottom_sweep<T, NZ>(L, 1, x, b);
        __syncthreads();
        bottom_ghost(L, x, b);
        __syncthreads();
    }
}

// The V cycle on the small levels in one workgroup (a workgroup barrier between steps instead of a
// kernel boundary).  All levels 1 .. nlev-1 of the group have NZ sigma levels; level 0 is solved
// densely.  The level table lives in device memory: each level's arguments are copied out of it once
// (no run-time indexing of a kernel-argument array, which would go through scratch).
template <typename T, int NZ>
__global__ __launch_bounds__(512) void k_mg_bottom(const BottomArgs<T> *__restrict__ A) {
    if (solve_done(A->skip)) return;
    const int top = A->nlev - 1, nu1 = A->nu1, nu2 = A->nu2;
    for (int l = top; l >= 1; l--) {
        const LvArgs<T> L = A->L[l];
        T *x = A->x[l], *r = A->r[l];
        const T *b = A->b[l];
        bottom_smooth<T, NZ>(L, nu1, x, b);
        bottom_residual<T, NZ>(L, x, b, r);
        __syncthreads();
        const RestrictGeo G = A->G[l];
        T *bc = A->b[l - 1], *xc = A->x[l - 1];
        const size_t nc = (size_t)G.nxgc * G.nygc * G.nzgc;
        for (size_t q = threadIdx.x; q < nc; q += blockDim.x) restrict_point(G, q, (const T *)r, bc, xc);
        __syncthreads();
    }
    {
        const int n0 = A->n0;
        const T *minv = A->minv, *b0 = A->b[0];
        T *x0 = A->x[0];
        for (int rr = threadIdx.x; rr < n0; rr += blockDim.x) {
            T s = T(0);
            for (int c = 0; c < n0; c++) s += minv[(size_t)rr * A->pitch0 + c] * b0[c];
            x0[rr] = s;
        }
    }
    __syncthreads();
    for (int l = 1; l <= top; l++) {
        const RestrictGeo G = A->G[l];
        const LvArgs<T> L = A->L[l];
        T *x = A->x[l];
        const T *b = A->b[l], *xc = A->x[l - 1];
        const size_t nf = (size_t)G.nxgf * G.nygf * G.nzgf;
        for (size_t q = threadIdx.x; q < nf; q += blockDim.x) prolong_point(G, q, xc, x);
        __syncthreads();
        bottom_smooth<T, NZ>(L, nu2, x, b);
    }
}


}  // namespace ow3d
