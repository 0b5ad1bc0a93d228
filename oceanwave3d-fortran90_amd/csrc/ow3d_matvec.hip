// Dispatch of the high-order sigma-Laplace matvec kernels (k_matvec3c / k_matvec3p / k_matvec2p) over the
// column depth and the stencil half width.
#include "ow3d_internal.hpp"

// ---------------------------------------------------------------------------------------------
// Launch dispatch
// ---------------------------------------------------------------------------------------------
namespace ow3d_host {

namespace {

template <int NZ, int A, int TY>
constexpr size_t mv3_lds() {
    constexpr int R = A + 1, RA = 2 * A + 1;
    return ((size_t)(16 + 2 * R) * (TY + 2 * R) * NZ + 2 * 16 * TY * NZ + 7 * 16 * TY + (16 + 2) * 2 * RA + (TY + 2) * 2 * RA +
            NZ * 2 * RA + NZ) * 8;
}
template <int NZ, int A, int TY>
constexpr size_t mv3c_lds() {
    constexpr int R = A + 1, RA = 2 * A + 1;
    return ((size_t)NZ * (((16 + 2 * R) * (TY + 2 * R)) | 1) + (18 + TY + 2) * 2 * RA) * 8;
}
template <int NZ, int A>
int launch_mv3p(const ow3d_tank *t, const MvArgs &a, const double *phi, double *y, hipStream_t s) {
    constexpr int TY = mv3_lds<NZ, A, 8>() <= 65536 ? 8 : 4;
    static_assert(mv3_lds<NZ, A, TY>() <= 65536, "matvec tile exceeds 64 KiB of LDS");
    const int nb = ((t->g.nxg + 15) / 16) * ((t->g.nyg + TY - 1) / TY);
    hipLaunchKernelGGL((k_matvec3p<NZ, A, TY, 256>), dim3(nb), dim3(256), 0, s, a, phi, y);
    return 0;
}
template <int NZ, int A, int TY>
int launch_mv3c_ty(const ow3d_tank *t, const MvArgs &a, const double *phi, double *y, hipStream_t s) {
    static_assert(mv3c_lds<NZ, A, TY>() <= 65536, "matvec tile exceeds 64 KiB of LDS");
    const int nb = ((t->g.nxg + 15) / 16) * ((t->g.nyg + TY - 1) / TY);
    hipLaunchKernelGGL((k_matvec3c<NZ, A, TY>), dim3(nb), dim3(16 * TY), 0, s, a, phi, y);
    return 0;
}
template <int NZ, int A, int TX, int TY>
constexpr size_t mv3s_lds() {
    constexpr int R = A + 1, RA = 2 * A + 1;
    return ((size_t)NZ * (((TX + 2 * R) * (TY + 2 * R)) | 1) + (TX + 2 + TY + 2) * 2 * RA + NZ * TX * TY) * 8;
}
template <int NZ, int A, int TX, int TY>
int launch_mv3s_t(const ow3d_tank *t, const MvArgs &a, const double *phi, double *y, hipStream_t s) {
    static_assert(mv3s_lds<NZ, A, TX, TY>() <= 80 * 1024, "split matvec tile exceeds half the LDS of a CU");
    const int nb = ((t->g.nxg + TX - 1) / TX) * ((t->g.nyg + TY - 1) / TY);
    hipLaunchKernelGGL((k_matvec3s<NZ, A, TX, TY>), dim3(nb), dim3(2 * TX * TY), 0, s, a, phi, y);
    return 0;
}
template <int NZ, int A>
int launch_mv3s(const ow3d_tank *t, const MvArgs &a, const double *phi, double *y, hipStream_t s) {
    return launch_mv3s_t<NZ, A, 16, 8>(t, a, phi, y, s);
}
// matvec variant (ow3d_tune): 0 column kernel with the widest tile that fits 64 KiB of LDS, 1 column kernel
// with 16 x 8 tiles, 2 point kernel (k_matvec3p), 3 column kernel with each column split over two waves
// (16 x 8 tiles)
template <int NZ, int A>
int launch_mv3(const ow3d_tank *t, const MvArgs &a, const double *phi, double *y, hipStream_t s) {
    constexpr int TY = mv3c_lds<NZ, A, 16>() <= 65536 ? 16 : 8;
    if (t->mv_variant == 0) return launch_mv3c_ty<NZ, A, TY>(t, a, phi, y, s);
    if (t->mv_variant == 1) return launch_mv3c_ty<NZ, A, 8>(t, a, phi, y, s);
    if (t->mv_variant == 3) return launch_mv3s<NZ, A>(t, a, phi, y, s);
    return launch_mv3p<NZ, A>(t, a, phi, y, s);
}

template <int NZ, int A>
int launch_mv2(const ow3d_tank *t, const MvArgs &a, const double *phi, double *y, hipStream_t s) {
    dim3 grid((t->g.nxg + 63) / 64);
    hipLaunchKernelGGL((k_matvec2p<NZ, A>), grid, dim3(256), 0, s, a, phi, y);
    return 0;
}

template <int NZ>
int dispatch_mv_a(const ow3d_tank *t, const MvArgs &a, const double *phi, double *y, hipStream_t s) {
    const bool d3 = t->g.ny > 1;
    switch (t->A) {
        case 1: return d3 ? launch_mv3<NZ, 1>(t, a, phi, y, s) : launch_mv2<NZ, 1>(t, a, phi, y, s);
        case 2: return d3 ? launch_mv3<NZ, 2>(t, a, phi, y, s) : launch_mv2<NZ, 2>(t, a, phi, y, s);
        case 3: return d3 ? launch_mv3<NZ, 3>(t, a, phi, y, s) : launch_mv2<NZ, 3>(t, a, phi, y, s);
        case 4: return d3 ? launch_mv3<NZ, 4>(t, a, phi, y, s) : launch_mv2<NZ, 4>(t, a, phi, y, s);
    }
    return -1;
}

}  // namespace

int launch_matvec(const ow3d_tank *t, const double *phi, double *y, hipStream_t s, const int *skip) {
    MvArgs a = t->mv;
    a.skip = skip;
    switch (t->g.nzg) {
        case 5: return dispatch_mv_a<5>(t, a, phi, y, s);
        case 6: return dispatch_mv_a<6>(t, a, phi, y, s);
        case 7: return dispatch_mv_a<7>(t, a, phi, y, s);
        case 8: return dispatch_mv_a<8>(t, a, phi, y, s);
        case 9: return dispatch_mv_a<9>(t, a, phi, y, s);
        case 10: return dispatch_mv_a<10>(t, a, phi, y, s);
        case 11: return dispatch_mv_a<11>(t, a, phi, y, s);
        case 12: return dispatch_mv_a<12>(t, a, phi, y, s);
        case 13: return dispatch_mv_a<13>(t, a, phi, y, s);
        case 14: return dispatch_mv_a<14>(t, a, phi, y, s);
    }
    return -1;
}

bool matvec_supported(int nzg, int A) {
    static const int nz_ok[] = {5, 6, 7, 8, 9, 10, 11, 12, 13, 14};
    bool ok = false;
    for (int v : nz_ok) ok |= (v == nzg);
    return ok && A >= 1 && A <= 4 && 2 * A + 1 <= nzg;
}

}  // namespace ow3d_host
