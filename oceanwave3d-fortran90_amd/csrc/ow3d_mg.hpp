// Multigrid preconditioner internals shared by the per-precision translation units ow3d_mg64.hip and
// ow3d_mg32.hip (level set-up, kernel dispatch over the column depth, the V cycle).
#pragma once
#include <thread>
#include "ow3d_internal.hpp"

namespace ow3d_host {
namespace mg {

// LDS bytes of the tiled residual (16 x TY columns) and of the fused sweep (32 x TY columns).
template <typename T, int NZ, int TY>
constexpr size_t residual_lds() {
    constexpr int WX = 20, WY = TY + 4, RS2 = 16 * TY + 4, LX = WX * WY > RS2 ? WX * WY : RS2;
    constexpr int C3 = CoefLds<WX, WY, NZ>::TOTAL, C2 = CoefLds<RS2, 1, NZ>::TOTAL;
    return ((size_t)LX * (NZ | 1) + (C3 > C2 ? C3 : C2)) * sizeof(T);
}
template <typename T, int NZ, int TX, int TY>
constexpr size_t rb_lds() {
    constexpr int NS = ((TX + 2) / 2) * (TY + 2);
    return ((size_t)(TX + 4) * (TY + 4) * (NZ | 1) + CoefLds<TX + 4, TY + 4, NZ>::TOTAL + 4 * NS * NZ + NS +
            (size_t)(TX + 2) * (TY + 2) * NZ) * sizeof(T);
}
// The same with the prolongation's coarse window and tables (MODE 3 and 4).
template <typename T, int NZ, int TX, int TY>
constexpr size_t rb_lds_pro() {
    using W = ProWindow<TX + 4, TY + 4>;
    return rb_lds<T, NZ, TX, TY>() + (size_t)W::CWX * W::CWY * (NZ | 1) * sizeof(T) + (size_t)(TX + TY + 8) * (sizeof(int2) + 3 * sizeof(T));
}

template <typename T, int NZ>
void mg_launch_nz(const LvArgs<T> &L, int what, int colour, T *x, const T *b, T *r, hipStream_t s) {
    if (what == 0) {
        // 16 x 8 tiles (128 threads): on the 513^2 x 10 hierarchy 57.6 us of residuals per FP32 V cycle
        // against 64.0 us with 16 x 16 tiles (more blocks per CU; the small levels gain most)
        constexpr int TY = 8;
        static_assert(residual_lds<T, NZ, TY>() <= 65536, "residual tile exceeds 64 KiB of LDS");
        dim3 grid = L.dim3 ? dim3((L.nxg + 15) / 16, (L.nyg + TY - 1) / TY) : dim3((L.nxg + 16 * TY - 1) / (16 * TY), 1);
        hipLaunchKernelGGL((k_mg_residual_t<T, NZ, TY>), grid, dim3(16 * TY), 0, s, L, (const T *)x, b, r);
    } else {
        const int nxi = L.nxg - 2 * L.gx, nyi = L.dim3 ? L.nyg - 2 * L.gy : 1;
        constexpr size_t lds = (size_t)(34 * 10 + 32 * 8) * (NZ | 1) * sizeof(T);
        if constexpr (lds <= 65536) {
            dim3 grid((nxi + 31) / 32, L.dim3 ? (nyi + 7) / 8 : 1);
            hipLaunchKernelGGL((k_mg_zebra_t<T, NZ>), grid, dim3(128), 0, s, L, colour, x, b);
        } else {
            const int half = (nxi + 1) / 2;
            hipLaunchKernelGGL((k_mg_zebra<T, NZ>), dim3((half * nyi + 255) / 256), dim3(256), 0, s, L, colour, x, b);
        }
    }
}

template <typename T>
int mg_launch(const LvArgs<T> &L, int what, int colour, T *x, const T *b, T *r, hipStream_t s) {
    switch (L.nzg) {
#define OW3D_NZ(N) case N: mg_launch_nz<T, N>(L, what, colour, x, b, r, s); return 0;
        OW3D_NZ(3) OW3D_NZ(4) OW3D_NZ(5) OW3D_NZ(6) OW3D_NZ(7) OW3D_NZ(8) OW3D_NZ(9) OW3D_NZ(10)
        OW3D_NZ(11) OW3D_NZ(12) OW3D_NZ(13) OW3D_NZ(14)
#undef OW3D_NZ
    }
    return -1;
}

// ---------------------------------------------------------------------------------------------
// Multigrid set-up and V cycle (MultiGridSolver.f90:5-192 structure with the zebra smoother)
// ---------------------------------------------------------------------------------------------
// Inverse of the coarsest level operator by LU with partial pivoting.  The operator is banded (kl rows
// below, ku above the diagonal in the sigma-fastest ordering), so elimination and substitution only
// visit the band (pivoting widens U to ku + kl); every skipped term is an exact zero, so the result is
// the one dense elimination gives.  The n columns of the inverse are solved on up to 16 host threads.
inline void dense_inverse(const std::vector<double> &A, int n, std::vector<double> &Ainv) {
    std::vector<double> M(A);
    int kl = 0, ku = 0;
    for (int r = 0; r < n; r++)
        for (int c = 0; c < n; c++)
            if (M[(size_t)r * n + c] != 0.0) { kl = std::max(kl, r - c); ku = std::max(ku, c - r); }
    const int kw = ku + kl;   // U's band after row interchanges
    std::vector<int> piv(n);
    for (int c = 0; c < n; c++) {
        const int rend = std::min(n, c + kl + 1), qend = std::min(n, c + kw + 1);
        int pr = c; double best = std::fabs(M[(size_t)c * n + c]);
        for (int r = c + 1; r < rend; r++) if (std::fabs(M[(size_t)r * n + c]) > best) { best = std::fabs(M[(size_t)r * n + c]); pr = r; }
        piv[c] = pr;
        if (pr != c) for (int q = 0; q < n; q++) std::swap(M[(size_t)c * n + q], M[(size_t)pr * n + q]);
        for (int r = c + 1; r < rend; r++) {
            double f = M[(size_t)r * n + c] / M[(size_t)c * n + c];
            M[(size_t)r * n + c] = f;
            for (int q = c + 1; q < qend; q++) M[(size_t)r * n + q] -= f * M[(size_t)c * n + q];
        }
    }
    Ainv.assign((size_t)n * n, 0.0);
    auto solve_cols = [&](int c0, int c1) {
        std::vector<double> b(n);
        for (int col = c0; col < c1; col++) {
            std::fill(b.begin(), b.end(), 0.0); b[col] = 1.0;
            for (int c = 0; c < n; c++) if (piv[c] != c) std::swap(b[c], b[piv[c]]);
            int first = 0;
            while (first < n && b[first] == 0.0) first++;
            for (int r = first; r < n; r++) {
                double s = b[r];
                for (int q = std::max(first, r - kl); q < r; q++) s -= M[(size_t)r * n + q] * b[q];
                b[r] = s;
            }
            for (int r = n - 1; r >= 0; r--) {
                double s = b[r];
                const int qend = std::min(n, r + kw + 1);
                for (int q = r + 1; q < qend; q++) s -= M[(size_t)r * n + q] * b[q];
                b[r] = s / M[(size_t)r * n + r];
            }
            for (int r = 0; r < n; r++) Ainv[(size_t)r * n + col] = b[r];
        }
    };
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 512 || nt == 1) { solve_cols(0, n); return; }
    std::vector<std::thread> th;
    for (int k = 0; k < nt; k++) th.emplace_back(solve_cols, (int)((long)n * k / nt), (int)((long)n * (k + 1) / nt));
    for (auto &x : th) x.join();
}

// Host evaluation of the level operator applied to a unit vector (coarsest level assembly).
// Same row semantics as k_mg_residual_t, in double.
inline void low_apply_host(const Grid &g, const std::vector<double> &x, std::vector<double> &y) {
    const int nxg = g.nxg, nyg = g.nyg, nzg = g.nzg, nxy = g.nxy;
    const bool d3 = g.ny > 1;
    auto X = [&](int k, int i, int j) { return x[((size_t)j * nxg + i) * nzg + k]; };
    auto coefs = [&](int p, int k, double &s2, double &s3, double &s4, double &s5) {
        double zs = k == 0 ? 0.0 : g.z[k], d = g.h[p], hx = g.hx[p], hy = d3 ? g.hy[p] : 0.0;
        double hs = g.hxx[p] + (d3 ? g.hyy[p] : 0.0);
        double s1 = zs;
        s2 = ((1.0 - s1) * hx - s1 * 0.0) / d;
        if (d3) s3 = -(2.0 * (1.0 - s1) * (hx * hx + hy * hy) / (d * d)) + 2.0 * (2.0 * s1 - 1.0) * (0.0 * hx + 0.0 * hy) / (d * d) + (1.0 - s1) * hs / d + 2.0 * s1 * 0.0 / (d * d) - s1 * 0.0 / d;
        else s3 = -(2.0 * (1.0 - s1) * (hx * hx) / (d * d)) + 2.0 * (2.0 * s1 - 1.0) * (0.0 * hx) / (d * d) + (1.0 - s1) * hs / d + 2.0 * s1 * 0.0 / (d * d) - s1 * 0.0 / d;
        s4 = d3 ? ((1.0 - s1) * hy - s1 * 0.0) / d : 0.0;
        s5 = 1.0 / d;
    };
    y.assign(g.n, 0.0);
    for (int p = 0; p < nxy; p++) {
        int i = p % nxg, j = p / nxg;
        double *yc = &y[(size_t)p * nzg];
        bool wx = g.gx && (i == 0 || i == nxg - 1), wy = g.gy && (j == 0 || j == nyg - 1);
        if (wx && wy) { for (int k = 0; k < nzg; k++) yc[k] = X(k, i, j); continue; }
        if (wx || wy) {
            yc[0] = X(0, i, j);
            if (wx) {
                int ib = i == 0 ? 1 : nxg - 2, ws = clampi(ib - 1, 0, nxg - 3);
                const double *S = &g.sx[(size_t)(ib * 2) * 3];
                for (int k = 1; k < nzg; k++) yc[k] = (S[0] * X(k, ws, j) + S[1] * X(k, ws + 1, j)) + S[2] * X(k, ws + 2, j);
            } else {
                int jb = j == 0 ? 1 : nyg - 2, ws = clampi(jb - 1, 0, nyg - 3);
                const double *S = &g.sy[(size_t)(jb * 2) * 3];
                for (int k = 1; k < nzg; k++) yc[k] = (S[0] * X(k, i, ws) + S[1] * X(k, i, ws + 1)) + S[2] * X(k, i, ws + 2);
            }
            continue;
        }
        int wsx = clampi(i - 1, 0, nxg - 3), wsy = d3 ? clampi(j - 1, 0, nyg - 3) : 0;
        const double *SX1 = &g.sx[(size_t)(i * 2) * 3], *SX2 = SX1 + 3;
        const double *SY1 = d3 ? &g.sy[(size_t)(j * 2) * 3] : nullptr, *SY2 = d3 ? SY1 + 3 : nullptr;
        std::vector<double> dxl(nzg), dyl(nzg, 0.0);
        for (int k = 0; k < nzg; k++) {
            dxl[k] = (SX1[0] * X(k, wsx, j) + SX1[1] * X(k, wsx + 1, j)) + SX1[2] * X(k, wsx + 2, j);
            if (d3) dyl[k] = (SY1[0] * X(k, i, wsy) + SY1[1] * X(k, i, wsy + 1)) + SY1[2] * X(k, i, wsy + 2);
        }
        int sh = 0; if (wsx == 0) sh = 1; if (wsx + 2 == nxg - 1) sh = -1;
        if (sh) { const double *S = &g.sx[(size_t)((i - sh) * 2) * 3]; dxl[0] = (S[0] * X(0, wsx + sh, j) + S[1] * X(0, wsx + sh + 1, j)) + S[2] * X(0, wsx + sh + 2, j); }
        if (d3) {
            sh = 0; if (wsy == 0) sh = 1; if (wsy + 2 == nyg - 1) sh = -1;
            if (sh) { const double *S = &g.sy[(size_t)((j - sh) * 2) * 3]; dyl[0] = (S[0] * X(0, i, wsy + sh) + S[1] * X(0, i, wsy + sh + 1)) + S[2] * X(0, i, wsy + sh + 2); }
        }
        const double *xc = &x[(size_t)p * nzg];
        yc[nzg - 1] = xc[nzg - 1];
        for (int k = 1; k < nzg - 1; k++) {
            const double *Z1 = &g.sz[(size_t)(k * 2) * 3], *Z2 = Z1 + 3;
            double dz = (Z1[0] * xc[k - 1] + Z1[1] * xc[k]) + Z1[2] * xc[k + 1];
            double dzz = (Z2[0] * xc[k - 1] + Z2[1] * xc[k]) + Z2[2] * xc[k + 1];
            double dxx = (SX2[0] * X(k, wsx, j) + SX2[1] * X(k, wsx + 1, j)) + SX2[2] * X(k, wsx + 2, j);
            double dyy = d3 ? (SY2[0] * X(k, i, wsy) + SY2[1] * X(k, i, wsy + 1)) + SY2[2] * X(k, i, wsy + 2) : 0.0;
            double dxz = (Z1[0] * dxl[k - 1] + Z1[1] * dxl[k]) + Z1[2] * dxl[k + 1];
            double dyz = (Z1[0] * dyl[k - 1] + Z1[1] * dyl[k]) + Z1[2] * dyl[k + 1];
            double s2, s3, s4, s5; coefs(p, k, s2, s3, s4, s5);
            yc[k] = dxx + dyy + s3 * dz + 2.0 * (s2 * dxz + s4 * dyz) + (s2 * s2 + s4 * s4 + s5 * s5) * dzz;
        }
        const double *Z1 = &g.sz[(size_t)(1 * 2) * 3];
        double dz = (Z1[0] * xc[0] + Z1[1] * xc[1]) + Z1[2] * xc[2];
        double s2, s3, s4, s5; coefs(p, 1, s2, s3, s4, s5);
        double hx = g.hx[p], hy = d3 ? g.hy[p] : 0.0;
        double dx1 = (SX1[0] * X(1, wsx, j) + SX1[1] * X(1, wsx + 1, j)) + SX1[2] * X(1, wsx + 2, j);
        double dy1 = d3 ? (SY1[0] * X(1, i, wsy) + SY1[1] * X(1, i, wsy + 1)) + SY1[2] * X(1, i, wsy + 2) : 0.0;
        yc[0] = (s5 + hx * s2 + hy * s4) * dz + hx * dx1 + hy * dy1;
    }
}

// True when the 3-point rows S[N][2][3] of nodes 2 .. N-3 are the centred rows of a uniform grid: all
// equal (to 1e-12) to the middle node's, with a zero centre weight in the first derivative.  c receives
// the first-derivative weight of the right neighbour, the second-derivative neighbour and centre weights.
inline bool uniform_rows(const std::vector<double> &S, int N, double *c) {
    if (N < 5) return false;
    const double *R = &S[(size_t)(N / 2) * 6];
    c[0] = R[2]; c[1] = R[3]; c[2] = R[4];
    const double t1 = 1e-12 * std::fabs(c[0]), t2 = 1e-12 * std::fabs(c[2]);
    for (int i = 2; i <= N - 3; i++) {
        const double *Q = &S[(size_t)i * 6];
        if (std::fabs(Q[0] + c[0]) > t1 || std::fabs(Q[1]) > t1 || std::fabs(Q[2] - c[0]) > t1 ||
            std::fabs(Q[3] - c[1]) > t2 || std::fabs(Q[4] - c[2]) > t2 || std::fabs(Q[5] - c[1]) > t2)
            return false;
    }
    return true;
}

template <typename T>
int mg_build(ow3d_tank *t, std::vector<std::unique_ptr<MgLevel<T>>> &mg) {
    hipStream_t s = t->s;
    mg.clear();
    for (int k = 0; k < t->nlev; k++) {
        auto L = std::make_unique<MgLevel<T>>();
        L->g = t->levels[k];
        const Grid &g = L->g;
        auto up = [&](DevBuf<T> &d, const std::vector<double> &src) -> hipError_t {
            std::vector<T> tmp(src.begin(), src.end());
            hipError_t e = d.alloc(tmp.size());
            if (e != hipSuccess) return e;
            e = hipMemcpy(d.p, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice);
            return e;
        };
        std::vector<double> sy = g.sy;
        if (g.ny == 1) sy.assign(6, 0.0);
        std::vector<double> zs(g.z); zs[0] = 0.0;
        std::vector<double> col((size_t)6 * g.nxy);
        for (int p = 0; p < g.nxy; p++) {
            const double d = g.h[p], hx = g.hx[p], hy = g.ny > 1 ? g.hy[p] : 0.0;
            const double hs = g.hxx[p] + (g.ny > 1 ? g.hyy[p] : 0.0);
            col[p] = 1.0 / d; col[g.nxy + p] = hx; col[2 * g.nxy + p] = hy;
            col[3 * g.nxy + p] = hx / d; col[4 * g.nxy + p] = hy / d;
            col[5 * g.nxy + p] = -(2.0 * (hx * hx + hy * hy) / (d * d)) + hs / d;
        }
        HIPCHK(up(L->sx, g.sx)); HIPCHK(up(L->sy, sy)); HIPCHK(up(L->sz, g.sz)); HIPCHK(up(L->zs, zs)); HIPCHK(up(L->col, col));
        // the coarsest level's b is padded with zeros to the inverse's row pitch (whole 16-byte loads)
        const size_t nb = k == 0 ? (size_t)minv_pitch((int)g.n) : g.n;
        HIPCHK(L->x.alloc(g.n)); HIPCHK(L->x2.alloc(g.n)); HIPCHK(L->b.alloc(nb)); HIPCHK(L->r.alloc(g.n));
        if (k == 0) HIPCHK(hipMemset(L->b.p, 0, nb * sizeof(T)));
        HIPCHK(L->xf.alloc(g.nxg)); HIPCHK(L->yf.alloc(g.nyg)); HIPCHK(L->zf.alloc(g.nzg));
        HIPCHK(hipMemcpy(L->xf.p, g.x.data(), g.nxg * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(L->yf.p, g.y.data(), g.nyg * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(L->zf.p, g.z.data(), g.nzg * 8, hipMemcpyHostToDevice));
        if (k == 0) {
            int n = (int)g.n;
            std::vector<double> Ainv;
            {   // host work only: the capture lock is released meanwhile (see create_lock)
                std::unique_lock<std::mutex> *cl = create_lock();
                if (cl) cl->unlock();
                std::vector<double> A((size_t)n * n), e(n, 0.0), colv;
                for (int c = 0; c < n; c++) {
                    e[c] = 1.0; low_apply_host(g, e, colv); e[c] = 0.0;
                    for (int r = 0; r < n; r++) A[(size_t)r * n + c] = colv[r];
                }
                dense_inverse(A, n, Ainv);
                if (cl) cl->lock();
            }
            // rows padded with zeros to a multiple of 8 elements, so every row starts 16-byte aligned
            const int P = minv_pitch(n);
            L->minv_pitch = P;
            if (t->cfg.mg_precision == 2) {   // FP16 storage of the coarsest level's inverse
                std::vector<__half> h((size_t)n * P, __float2half(0.0f));
                for (int r = 0; r < n; r++)
                    for (int c = 0; c < n; c++) h[(size_t)r * P + c] = __float2half((float)Ainv[(size_t)r * n + c]);
                HIPCHK(L->minv16.alloc(h.size()));
                HIPCHK(hipMemcpy(L->minv16.p, h.data(), h.size() * sizeof(__half), hipMemcpyHostToDevice));
            } else {
                std::vector<double> pad((size_t)n * P, 0.0);
                for (int r = 0; r < n; r++) std::copy(&Ainv[(size_t)r * n], &Ainv[(size_t)r * n] + n, &pad[(size_t)r * P]);
                HIPCHK(up(L->minv, pad));
            }
        }
        L->args.nxg = g.nxg; L->args.nyg = g.nyg; L->args.nzg = g.nzg; L->args.nxy = g.nxy;
        L->args.gx = g.gx; L->args.gy = g.gy; L->args.dim3 = g.ny > 1;
        L->args.sx = L->sx.p; L->args.sy = L->sy.p; L->args.sz = L->sz.p; L->args.zs = L->zs.p; L->args.col = L->col.p;
        L->args.skip = t->done();
        // a launch on a level of up to 2^20 points is a chain of memory latencies, not a stream of
        // bytes: its kernels issue their loads before testing skip (one wait instead of two), at the cost
        // of reading the level once more in launches that follow convergence
        L->args.spec = g.n <= (1u << 20) ? 1 : 0;
        {   // uniform horizontal grid: the interior columns' centred rows are all equal to the middle one's
            double c[6] = {0, 0, 0, 0, 0, 0};
            const bool ux = uniform_rows(g.sx, g.nxg, c), uy = g.ny == 1 || uniform_rows(g.sy, g.nyg, c + 3);
            L->args.uni = ux && uy;
            L->args.cx1 = (T)c[0]; L->args.cx2 = (T)c[1]; L->args.cxc = (T)c[2];
            L->args.cy1 = (T)c[3]; L->args.cy2 = (T)c[4]; L->args.cyc = (T)c[5];
        }
        mg.push_back(std::move(L));
    }
    // interpolation tables of the prolongations (see MgLevel::ptab): the weights are the ones prolong_at
    // forms per point, (T) of the double coordinate differences and T(1) / (their sum) in T
    for (int k = 1; k < t->nlev; k++) {
        MgLevel<T> &F = *mg[k];
        const Grid &f = F.g, &c = mg[k - 1]->g;
        if (!(c.xc && c.yc && !c.zc && f.ny > 1)) continue;
        auto table = [](int n, int g, int nin, const std::vector<double> &x, std::vector<int2> &idx, std::vector<T> &w) {
            idx.assign(n, make_int2(-1, -1));
            w.assign((size_t)3 * n, T(0));
            for (int i = 0; i < n; i++) {
                if (i < g || i > nin + g - 1) continue;
                const int ie = i - g;
                if (ie % 2 == 0) {
                    idx[i] = make_int2(g + ie / 2, g + ie / 2);
                    w[3 * i] = T(1); w[3 * i + 1] = T(0); w[3 * i + 2] = T(1);
                } else {
                    const int cl = g + (ie - 1) / 2;
                    const T dlo = (T)(x[i] - x[i - 1]), dhi = (T)(x[i + 1] - x[i]);
                    idx[i] = make_int2(cl, cl + 1);
                    w[3 * i] = dhi; w[3 * i + 1] = dlo; w[3 * i + 2] = T(1) / (dlo + dhi);
                }
            }
        };
        std::vector<int2> xi, yi;
        std::vector<T> xw, yw;
        table(f.nxg, f.gx, f.nx, f.x, xi, xw);
        table(f.nyg, f.gy, f.ny, f.y, yi, yw);
        HIPCHK(F.pxi.alloc(xi.size())); HIPCHK(F.pyi.alloc(yi.size()));
        HIPCHK(F.pxw.alloc(xw.size())); HIPCHK(F.pyw.alloc(yw.size()));
        HIPCHK(hipMemcpy(F.pxi.p, xi.data(), xi.size() * sizeof(int2), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(F.pyi.p, yi.data(), yi.size() * sizeof(int2), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(F.pxw.p, xw.data(), xw.size() * sizeof(T), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(F.pyw.p, yw.data(), yw.size() * sizeof(T), hipMemcpyHostToDevice));
        // restriction weights per coarse column / row, as restrict_at forms them
        auto rtab = [](int nc, int g, int ncin, const std::vector<double> &x, std::vector<T> &w) {
            w.assign((size_t)4 * nc, T(0));
            for (int ic = 0; ic < nc; ic++) {
                const int icn = ic - g;
                if (icn < 0 || icn >= ncin) continue;
                const int i = g + 2 * icn;
                const T dlo = (T)(x[i] - x[i - 1]), dhi = (T)(x[i + 1] - x[i]);
                w[4 * ic] = dlo; w[4 * ic + 1] = dlo + dhi; w[4 * ic + 2] = dhi; w[4 * ic + 3] = T(0.5) * (T(1) / (dlo + dhi));
            }
        };
        std::vector<T> rx, ry;
        rtab(c.nxg, c.gx, c.nx, f.x, rx);
        rtab(c.nyg, c.gy, c.ny, f.y, ry);
        // the y smoothing sums (d_lo, d_hi + d_lo, d_hi) and scales by 1 / (d_lo + d_hi): same values
        for (int jc = 0; jc < c.nyg; jc++) ry[4 * jc + 1] = ry[4 * jc + 2] + ry[4 * jc];
        HIPCHK(F.rxw.alloc(rx.size())); HIPCHK(F.ryw.alloc(ry.size()));
        HIPCHK(hipMemcpy(F.rxw.p, rx.data(), rx.size() * sizeof(T), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(F.ryw.p, ry.data(), ry.size() * sizeof(T), hipMemcpyHostToDevice));
        F.ptab = true;
    }
    (void)s;
    return 0;
}

inline RestrictGeo geo(const Grid &f, const Grid &c, const double *xf, const double *yf, const double *zf, const int *skip) {
    RestrictGeo G{};
    G.nxgf = f.nxg; G.nygf = f.nyg; G.nzgf = f.nzg; G.nxf = f.nx; G.nyf = f.ny; G.nzf = f.nz; G.gx = f.gx; G.gy = f.gy;
    G.nxgc = c.nxg; G.nygc = c.nyg; G.nzgc = c.nzg; G.nxc = c.nx; G.nyc = c.ny; G.nzc = c.nz;
    G.xc = c.xc; G.yc = c.yc; G.zc = c.zc;
    G.xf = xf; G.yf = yf; G.zf = zf;
    G.skip = skip;
    return G;
}

template <typename T>
int fill_bottom(ow3d_tank *t, std::vector<std::unique_ptr<MgLevel<T>>> &mg, BottomArgs<T> &A, DevBuf<BottomArgs<T>> &dA) {
    A = BottomArgs<T>{};
    if (t->bot_top < 0) return 0;
    A.nlev = t->bot_top + 1;
    A.nu1 = t->cfg.nu1; A.nu2 = t->cfg.nu2;
    A.n0 = (int)mg[0]->g.n;
    A.pitch0 = mg[0]->minv_pitch;
    A.minv = mg[0]->minv.p;
    for (int k = 0; k <= t->bot_top; k++) {
        MgLevel<T> &L = *mg[k];
        A.L[k] = L.args; A.x[k] = L.x.p; A.b[k] = L.b.p; A.r[k] = L.r.p;
        if (k >= 1) A.G[k] = geo(L.g, mg[k - 1]->g, L.xf.p, L.yf.p, L.zf.p, t->done());
    }
    A.skip = t->done();
    HIPCHK(dA.alloc(1));
    HIPCHK(hipMemcpy(dA.p, &A, sizeof(A), hipMemcpyHostToDevice));
    return 0;
}

template <typename T>
void ghost_pass(MgLevel<T> &L, T *x, hipStream_t s) {
    const Grid &g = L.g;
    const int nb = (g.ny > 1 ? 2 * g.nxg + 2 * (g.nyg - 2) : 2) * g.nzg;
    hipLaunchKernelGGL((k_mg_ghost_p<T>), dim3((nb + 255) / 256), dim3(256), 0, s, L.args, x, (const T *)L.b.p);
}

// Restriction and prolongation: one grid row of columns per blockIdx.y with the level count a template
// argument (no runtime division per point) for nzg 3..14, else the flat kernels.
template <typename T>
void restrict_launch(const RestrictGeo &G, const T *rf, T *rc, T *xz, hipStream_t s) {
    switch (G.nzgc) {
#define OW3D_NZ(N) case N: hipLaunchKernelGGL((k_mg_restrict_t<T, N>), dim3((G.nxgc * N + 255) / 256, G.nygc), dim3(256), 0, s, G, rf, rc, xz); return;
        OW3D_NZ(3) OW3D_NZ(4) OW3D_NZ(5) OW3D_NZ(6) OW3D_NZ(7) OW3D_NZ(8) OW3D_NZ(9) OW3D_NZ(10)
        OW3D_NZ(11) OW3D_NZ(12) OW3D_NZ(13) OW3D_NZ(14)
#undef OW3D_NZ
    }
    hipLaunchKernelGGL((k_mg_restrict<T>), dim3(grid_blocks((size_t)G.nxgc * G.nygc * G.nzgc)), dim3(256), 0, s, G, rf, rc, xz);
}
// Restriction from fine level F with its weight tables (F.ptab).
template <typename T>
void restrict_tab_launch(const RestrictGeo &G, const MgLevel<T> &F, const T *rf, T *rc, T *xz, hipStream_t s) {
    switch (G.nzgc) {
#define OW3D_NZ(N) case N: hipLaunchKernelGGL((k_mg_restrict_tab<T, N>), dim3((G.nxgc * N + 255) / 256, G.nygc), dim3(256), 0, s, G, \
                                          (const T *)F.rxw.p, (const T *)F.ryw.p, rf, rc, xz); return;
        OW3D_NZ(3) OW3D_NZ(4) OW3D_NZ(5) OW3D_NZ(6) OW3D_NZ(7) OW3D_NZ(8) OW3D_NZ(9) OW3D_NZ(10)
        OW3D_NZ(11) OW3D_NZ(12) OW3D_NZ(13) OW3D_NZ(14)
#undef OW3D_NZ
    }
    hipLaunchKernelGGL((k_mg_restrict<T>), dim3(grid_blocks((size_t)G.nxgc * G.nygc * G.nzgc)), dim3(256), 0, s, G, rf, rc, xz);
}
// Residual of level F and its restriction in one launch (k_mg_residual_restrict, F.ptab levels), for the
// column depths whose tile fits 64 KiB of LDS; false otherwise (the caller then launches the two).  Coarse
// tiles of 8 x 7 columns: 17 x 15 fine columns, one per thread of 256.
template <typename T, int NZ>
bool resrestrict_nz(const MgLevel<T> &F, const RestrictGeo &G, const T *x, T *rc, T *xz, hipStream_t s) {
    constexpr int CX = 8, CY = 7;
    if constexpr (ResRestrictLds<NZ, CX, CY>::TOTAL * sizeof(T) <= 65536) {
        hipLaunchKernelGGL((k_mg_residual_restrict<T, NZ, CX, CY, 256>), dim3((G.nxgc + CX - 1) / CX, (G.nygc + CY - 1) / CY), dim3(256), 0, s,
                           F.args, G, (const T *)F.rxw.p, (const T *)F.ryw.p, x, (const T *)F.b.p, rc, xz);
        return true;
    }
    return false;
}
template <typename T>
bool resrestrict_launch(const MgLevel<T> &F, const RestrictGeo &G, const T *x, T *rc, T *xz, hipStream_t s) {
    if (!F.ptab || !F.args.dim3 || G.gx != 1 || G.gy != 1 || G.nzgc != F.g.nzg) return false;
    switch (F.g.nzg) {
#define OW3D_NZ(N) case N: return resrestrict_nz<T, N>(F, G, x, rc, xz, s);
        OW3D_NZ(3) OW3D_NZ(4) OW3D_NZ(5) OW3D_NZ(6) OW3D_NZ(7) OW3D_NZ(8) OW3D_NZ(9) OW3D_NZ(10)
        OW3D_NZ(11) OW3D_NZ(12) OW3D_NZ(13) OW3D_NZ(14)
#undef OW3D_NZ
    }
    return false;
}
// Prolongation onto fine level F from its interpolation tables (F.ptab), else prolong_launch.
template <typename T>
void prolong_tab_launch(const RestrictGeo &G, const MgLevel<T> &F, const T *ec, T *xf, hipStream_t s) {
    switch (G.nzgf) {
#define OW3D_NZ(N) case N: hipLaunchKernelGGL((k_mg_prolong_tab<T, N>), dim3((G.nxgf * N + 255) / 256, G.nygf), dim3(256), 0, s, G, \
                                          (const int2 *)F.pxi.p, (const T *)F.pxw.p, (const int2 *)F.pyi.p, (const T *)F.pyw.p, ec, xf); return;
        OW3D_NZ(3) OW3D_NZ(4) OW3D_NZ(5) OW3D_NZ(6) OW3D_NZ(7) OW3D_NZ(8) OW3D_NZ(9) OW3D_NZ(10)
        OW3D_NZ(11) OW3D_NZ(12) OW3D_NZ(13) OW3D_NZ(14)
#undef OW3D_NZ
    }
    hipLaunchKernelGGL((k_mg_prolong_add<T>), dim3(grid_blocks((size_t)G.nxgf * G.nygf * G.nzgf)), dim3(256), 0, s, G, ec, xf);
}
template <typename T>
void prolong_launch(const RestrictGeo &G, const T *ec, T *xf, hipStream_t s) {
    switch (G.nzgf) {
#define OW3D_NZ(N) case N: hipLaunchKernelGGL((k_mg_prolong_add_t<T, N>), dim3((G.nxgf * N + 255) / 256, G.nygf), dim3(256), 0, s, G, ec, xf); return;
        OW3D_NZ(3) OW3D_NZ(4) OW3D_NZ(5) OW3D_NZ(6) OW3D_NZ(7) OW3D_NZ(8) OW3D_NZ(9) OW3D_NZ(10)
        OW3D_NZ(11) OW3D_NZ(12) OW3D_NZ(13) OW3D_NZ(14)
#undef OW3D_NZ
    }
    hipLaunchKernelGGL((k_mg_prolong_add<T>), dim3(grid_blocks((size_t)G.nxgf * G.nygf * G.nzgf)), dim3(256), 0, s, G, ec, xf);
}

// Fused sweep on TX x TY tiles with NT threads (the launch grid the kernel's tile arithmetic assumes).
// With pro (not NULL) the sweep first adds the prolongated coarse correction; false (nothing launched)
// when that tile does not fit 64 KiB of LDS.
template <typename T, int NZ, int TX, int TY, int NT>
bool rb_tiles(const LvArgs<T> &L, const T *xin, T *xout, const T *b, hipStream_t s, const double *r0, T *bout, double *uout,
              const ProArgs<T> *pro) {
    static_assert(rb_lds<T, NZ, TX, TY>() <= 65536, "zebra sweep tile exceeds 64 KiB of LDS");
    const int nxi = L.nxg - 2 * L.gx, nyi = L.dim3 ? L.nyg - 2 * L.gy : 1;
    const int offx = (nxi % TX == 1) ? 1 : 0, offy = (L.dim3 && nyi % TY == 1) ? 1 : 0;
    const dim3 grid((nxi + offx + TX - 1) / TX, L.dim3 ? (nyi + offy + TY - 1) / TY : 1);
    const ProArgs<T> P = pro ? *pro : ProArgs<T>{};
    if (pro) {
        if constexpr (rb_lds_pro<T, NZ, TX, TY>() <= 65536) {
            if (uout) hipLaunchKernelGGL((k_mg_zebra_rb<T, NZ, TX, TY, NT, 4>), grid, dim3(NT), 0, s, L, xin, xout, b, r0, bout, uout, P);
            else hipLaunchKernelGGL((k_mg_zebra_rb<T, NZ, TX, TY, NT, 3>), grid, dim3(NT), 0, s, L, xin, xout, b, r0, bout, uout, P);
            return true;
        }
        return false;
    }
    if (r0) hipLaunchKernelGGL((k_mg_zebra_rb<T, NZ, TX, TY, NT, 1>), grid, dim3(NT), 0, s, L, xin, xout, b, r0, bout, uout, P);
    else if (uout) hipLaunchKernelGGL((k_mg_zebra_rb<T, NZ, TX, TY, NT, 2>), grid, dim3(NT), 0, s, L, xin, xout, b, r0, bout, uout, P);
    else hipLaunchKernelGGL((k_mg_zebra_rb<T, NZ, TX, TY, NT>), grid, dim3(NT), 0, s, L, xin, xout, b, r0, bout, uout, P);
    return true;
}

// Fused red-black sweep from xin into xout (out of place: a block's staged neighbours must not change
// under it while other blocks write their results).
template <typename T, int NZ>
bool rb_launch_nz(const LvArgs<T> &L, const T *xin, T *xout, const T *b, int threads, hipStream_t s,
                  const double *r0, T *bout, double *uout, const ProArgs<T> *pro) {
    // the widest tile (32, 16 or 8 columns, 4 rows) whose staged data fits in 56 KiB of LDS
    constexpr int TY = 4, TX = rb_lds<T, NZ, 32, TY>() <= 57344 ? 32 : (rb_lds<T, NZ, 16, TY>() <= 57344 ? 16 : 8);
    const int nxi = L.nxg - 2 * L.gx, nyi = L.dim3 ? L.nyg - 2 * L.gy : 1;
    const long tiles = (long)((nxi + TX - 1) / TX) * (L.dim3 ? (nyi + TY - 1) / TY : 1);
    // FP32 3D levels with many tiles (measured at 513^2 x 10 and 257^2 x 10, kernel medians per sweep):
    // 16 x 8 tiles of 512 threads on the largest (42.7 us against 49.0 us for 32 x 4 tiles of 1024
    // threads: with twice the blocks per CU one block's loads overlap another's rows), 32 x 4 tiles of
    // 512 threads on the next (20.4 against 22.2 us); smaller levels keep one 1024-thread block per tile
    if constexpr (sizeof(T) == 4 && TX == 32 && rb_lds<T, NZ, 16, 8>() <= 65536) {
        if (L.dim3 && threads != 256 && tiles >= 1500) return rb_tiles<T, NZ, 16, 8, 512>(L, xin, xout, b, s, r0, bout, uout, pro);
        if (L.dim3 && threads != 256 && tiles >= 400) return rb_tiles<T, NZ, 32, 4, 512>(L, xin, xout, b, s, r0, bout, uout, pro);
    }
    // levels with fewer tiles than CUs are latency-bound: give each tile 1024 threads (about one row of
    // the line systems per thread); larger levels use 256-thread blocks for occupancy
    const bool big = threads == 1024 || (threads == 0 && tiles < 256);
    if (big) return rb_tiles<T, NZ, TX, TY, 1024>(L, xin, xout, b, s, r0, bout, uout, pro);
    return rb_tiles<T, NZ, TX, TY, 256>(L, xin, xout, b, s, r0, bout, uout, pro);
}

// Fused sweep xin -> xout; with r0 (not NULL) the first sweep of an application: x = 0 and b = (T) r0,
// written to bout (the level's b); with uout (not NULL) the last: the result goes to uout in FP64 and
// xout is not written; with pro (not NULL) the prolongated coarse correction is added to xin's values
// first.  Returns 0 when launched, 1 when the prolongating form does not fit (nothing launched), -1 for
// a column depth outside 3..14.
template <typename T>
int rb_launch(const LvArgs<T> &L, const T *xin, T *xout, const T *b, int threads, hipStream_t s,
              const double *r0 = nullptr, T *bout = nullptr, double *uout = nullptr, const ProArgs<T> *pro = nullptr) {
    switch (L.nzg) {
#define OW3D_NZ(N) case N: return rb_launch_nz<T, N>(L, xin, xout, b, threads, s, r0, bout, uout, pro) ? 0 : 1;
        OW3D_NZ(3) OW3D_NZ(4) OW3D_NZ(5) OW3D_NZ(6) OW3D_NZ(7) OW3D_NZ(8) OW3D_NZ(9) OW3D_NZ(10)
        OW3D_NZ(11) OW3D_NZ(12) OW3D_NZ(13) OW3D_NZ(14)
#undef OW3D_NZ
    }
    return -1;
}

// ghost pass, nu x (zebra sweep, ghost pass) on the level's current iterate `cur` (L.x or L.x2; a fused
// sweep moves it to the other buffer).
// r0 (fused sweeps, nu >= 1 only): the level's x is zero and its b is (T) r0, both still to be set; the
// first sweep does it.  uout (fused sweeps, nu >= 1 only): the last sweep writes its result there in FP64
// (cur is then stale).
// pro (fused sweeps, nu >= 1 only): the first sweep adds the prolongated coarse correction to cur's values;
// returns 1 with nothing launched when that sweep form does not fit.
template <typename T>
int smooth(ow3d_tank *t, MgLevel<T> &L, int nu, T *&cur, hipStream_t s, const double *r0 = nullptr,
           double *uout = nullptr, const ProArgs<T> *pro = nullptr) {
    if (t->fused_rb) {   // each launch is (ghost pass, red, black, ghost pass)
        if (nu == 0) ghost_pass(L, cur, s);
        for (int it = 0; it < nu; it++) {
            T *nxt = (cur == L.x.p) ? L.x2.p : L.x.p;
            const int rc = rb_launch<T>(L.args, cur, nxt, L.b.p, t->rb_threads, s, it == 0 ? r0 : nullptr, L.b.p,
                                        it == nu - 1 ? uout : nullptr, it == 0 ? pro : nullptr);
            if (rc) return rc;
            cur = nxt;
        }
        return 0;
    }
    ghost_pass(L, cur, s);
    for (int it = 0; it < nu; it++) {
        if (mg_launch<T>(L.args, 1, 0, cur, L.b.p, nullptr, s)) return -1;
        if (mg_launch<T>(L.args, 1, 1, cur, L.b.p, nullptr, s)) return -1;
        ghost_pass(L, cur, s);
    }
    return 0;
}

template <typename T> const BottomArgs<T> *bottom_args(const ow3d_tank *t);
template <> inline const BottomArgs<double> *bottom_args<double>(const ow3d_tank *t) { return t->bot64d.p; }
template <> inline const BottomArgs<float> *bottom_args<float>(const ow3d_tank *t) { return t->bot32d.p; }

// V cycle on level k from its current iterate `cur` (on entry L.x for every level below the finest);
// on return `cur` points at the buffer holding the result.
template <typename T>
int vcycle(ow3d_tank *t, std::vector<std::unique_ptr<MgLevel<T>>> &mg, int k, T *&cur, hipStream_t s,
           const double *r0 = nullptr, double *uout = nullptr) {
    MgLevel<T> &L = *mg[k];
    if (k == t->bot_top) {   // in place on L.x
        switch (L.g.nzg) {
#define OW3D_NZ(N) case N: hipLaunchKernelGGL((k_mg_bottom<T, N>), dim3(1), dim3(512), 0, s, bottom_args<T>(t)); break;
            OW3D_NZ(3) OW3D_NZ(4) OW3D_NZ(5) OW3D_NZ(6) OW3D_NZ(7) OW3D_NZ(8) OW3D_NZ(9) OW3D_NZ(10)
            OW3D_NZ(11) OW3D_NZ(12) OW3D_NZ(13) OW3D_NZ(14)
#undef OW3D_NZ
            default: err() = "bottom level depth outside 3..14"; return -1;
        }
        cur = L.x.p;
        return 0;
    }
    if (k == 0) {
        if (L.minv16.p)
            hipLaunchKernelGGL((k_mg_coarse<T, __half>), dim3(((int)L.g.n + 3) / 4), dim3(256), 0, s, (int)L.g.n, L.minv_pitch,
                               (const __half *)L.minv16.p, (const T *)L.b.p, L.x.p, t->done());
        else
            hipLaunchKernelGGL((k_mg_coarse<T, T>), dim3(((int)L.g.n + 3) / 4), dim3(256), 0, s, (int)L.g.n, L.minv_pitch,
                               (const T *)L.minv.p, (const T *)L.b.p, L.x.p, t->done());
        cur = L.x.p;
        return 0;
    }
    MgLevel<T> &C = *mg[k - 1];
    if (smooth(t, L, t->cfg.nu1, cur, s, r0)) return -1;
    RestrictGeo G = geo(L.g, C.g, L.xf.p, L.yf.p, L.zf.p, t->done());
    G.spec = L.args.spec;
    if (!resrestrict_launch<T>(L, G, cur, C.b.p, C.x.p, s)) {
        if (mg_launch<T>(L.args, 0, 0, cur, L.b.p, L.r.p, s)) return -1;
        if (L.ptab) restrict_tab_launch<T>(G, L, (const T *)L.r.p, C.b.p, C.x.p, s);
        else restrict_launch<T>(G, (const T *)L.r.p, C.b.p, C.x.p, s);
    }
    T *ccur = C.x.p;
    if (vcycle(t, mg, k - 1, ccur, s)) return -1;
    // with fused sweeps the first post-smoothing sweep adds the prolongated correction itself; FP64 levels
    // above 2^20 points keep the separate prolongation (at 515^2 x 10 the coarse window takes the FP64
    // sweep tile from 50 to 57 KiB of LDS, two blocks per CU instead of three: 0.55 ms per RK4 step slower)
    if (L.ptab && t->fused_rb && t->cfg.nu2 >= 1 && L.args.dim3 && (sizeof(T) == 4 || L.args.spec)) {
        const ProArgs<T> P{(const T *)ccur, L.pxi.p, L.pyi.p, L.pxw.p, L.pyw.p, C.g.nxg};
        const int rc = smooth(t, L, t->cfg.nu2, cur, s, nullptr, uout, &P);
        if (rc <= 0) return rc;
    }
    if (L.ptab) prolong_tab_launch<T>(G, L, (const T *)ccur, cur, s);
    else prolong_launch<T>(G, (const T *)ccur, cur, s);
    if (smooth(t, L, t->cfg.nu2, cur, s, nullptr, uout)) return -1;
    return 0;
}

// The preconditioner body in precision T: input r (FP64) -> finest-level right-hand side, V cycles from
// x = 0.  Only kernel launches (no memset / memcpy nodes), so the sequence can be captured in a hipGraph.
// Returns the buffer holding the result in *out.
// With u (not NULL) and post-smoothing on the finest level, the last sweep writes the result to u in FP64
// and *out is set to NULL; otherwise *out is the finest-level buffer holding it.
template <typename T>
int precond_body(ow3d_tank *t, std::vector<std::unique_ptr<MgLevel<T>>> &mg, const double *r, const void **out, hipStream_t s,
                 double *u) {
    const size_t n = t->g.n;
    const int top = t->nlev - 1;
    MgLevel<T> &F = *mg[top];
    // with fused sweeps and pre-smoothing above the coarsest level, the first sweep sets b = (T) r from
    // x = 0 itself; otherwise one pass does it first
    const bool fused = t->fused_rb && top >= 1 && top != t->bot_top;
    const bool fuse_in = fused && t->cfg.nu1 >= 1, fuse_out = fused && u && t->cfg.nu2 >= 1;
    if (!fuse_in)
        hipLaunchKernelGGL((k_precond_in<T>), dim3(grid_blocks(n)), dim3(256), 0, s, n, r, F.b.p, F.x.p, t->done());
    T *cur = F.x.p;
    const int cycles = t->nlev == 1 ? 1 : std::max(1, t->cfg.maxit);
    for (int it = 0; it < cycles; it++) {
        if (vcycle<T>(t, mg, top, cur, s, (fuse_in && it == 0) ? r : nullptr, (fuse_out && it == cycles - 1) ? u : nullptr)) return -1;
    }
    *out = fuse_out ? nullptr : cur;
    return 0;
}

}  // namespace mg
}  // namespace ow3d_host
