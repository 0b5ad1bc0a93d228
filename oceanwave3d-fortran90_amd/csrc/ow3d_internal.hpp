// Declarations shared by the host translation units of libow3d (ow3d.hip: C ABI, Krylov solvers and
// RK4; ow3d_matvec.hip: high-order operator dispatch; ow3d_mg.hip: multigrid set-up and V cycle).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/ow3d.h"
#include "ow3d_grid.hpp"
#include "ow3d_kernels.hpp"

namespace ow3d_host {
using namespace ow3d;

// message of the last failure on this host thread (ow3d_last_error)
std::string &err();

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                             \
            ow3d_host::err() = std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr; \
            return -1;                                                                      \
        }                                                                                   \
    } while (0)

constexpr int MAXK = 32;     // Krylov basis limit (GMRESmaxiterations <= MAXK - 1)
constexpr int RED_BLOCKS = 2048;

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    ~DevBuf() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t count) {
        n = count;
        return hipMalloc((void **)&p, (count ? count : 1) * sizeof(T));
    }
    hipError_t upload(const T *h, size_t count, hipStream_t s) {
        return hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s);
    }
};

template <typename T>
struct MgLevel {
    Grid g;
    DevBuf<T> sx, sy, sz, zs, col, x, x2, b, r, minv;   // x2: second iterate buffer of out-of-place sweeps
    DevBuf<double> xf, yf, zf;   // level coordinates (restriction / prolongation weights)
    LvArgs<T> args{};
};

inline int grid_blocks(size_t n, int bs = 256) {
    size_t b = (n + bs - 1) / bs;
    return (int)std::min<size_t>(std::max<size_t>(b, 1), 65535 * 16);
}

}  // namespace ow3d_host

using namespace ow3d;
using namespace ow3d_host;

struct ow3d_tank {
    ow3d_config cfg{};
    Grid g;
    hipStream_t s = nullptr;
    int A = 3;
    // high-order operator
    DevBuf<double> sx, sy, sz, zs, bath, h, cf;
    MvArgs mv{};
    // surface state
    DevBuf<double> E, P, W, Ex, Exx, Ey, Eyy, Px, Py, Et, Pt, Wt, KE, KP;
    // Laplace solve
    DevBuf<double> phi, rhs, last0, last1, w, z, V, part, red, hv;
    DevBuf<double> io_a, io_b;   // staging for host-pointer entry points
    // relaxation
    std::vector<Zone> zones;
    std::vector<double> gam;
    DevBuf<Zone> dzones;
    DevBuf<double> dgam;
    // multigrid
    int nlev = 0;
    std::vector<Grid> levels;
    std::vector<std::unique_ptr<MgLevel<double>>> mg64;
    std::vector<std::unique_ptr<MgLevel<float>>> mg32;
    int bot_top = -1;                 // finest level handled by the one-workgroup bottom kernel (-1: none)
    bool fused_rb = true;             // one launch per zebra sweep (red and black halves) on the upper levels
    BottomArgs<double> bot64{};
    BottomArgs<float> bot32{};
    // statistics
    long long it_total = 0, it_solves = 0;
    int it_min = -1, it_max = 0;
    long tstep = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool use_graph = true;
    hipGraphExec_t pg_exec = nullptr;   // captured preconditioner (V cycles)
    const void *mg_out = nullptr;       // finest-level buffer holding the preconditioner result
    ~ow3d_tank() {
        if (pg_exec) (void)hipGraphExecDestroy(pg_exec);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (s) (void)hipStreamDestroy(s);
    }
};

namespace ow3d_host {
// high-order operator (ow3d_matvec.hip)
int launch_matvec(const ow3d_tank *t, const double *phi, double *y, hipStream_t s);
bool matvec_supported(int nzg, int A);
// multigrid (ow3d_mg.hip): hierarchy set-up for t->cfg and u = M^-1 r on device vectors
int mg_setup(ow3d_tank *t);
int precond(ow3d_tank *t, const double *r, double *u, hipStream_t s);
// per-precision parts (ow3d_mg64.hip, ow3d_mg32.hip)
int mg_build64(ow3d_tank *t);
int mg_build32(ow3d_tank *t);
void mg_fill_bottom64(ow3d_tank *t);
void mg_fill_bottom32(ow3d_tank *t);
int precond_body64(ow3d_tank *t, const double *r, const void **out, hipStream_t s);
int precond_body32(ow3d_tank *t, const double *r, const void **out, hipStream_t s);
}  // namespace ow3d_host
