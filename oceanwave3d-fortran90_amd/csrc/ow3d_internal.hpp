// Declarations shared by the host translation units of libow3d (ow3d.hip: C ABI, Krylov solvers and
// RK4; ow3d_matvec.hip: high-order operator dispatch; ow3d_mg.hip: multigrid set-up and V cycle).
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <map>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
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

constexpr int MAXK = KRYLOV_MAXK;   // Krylov basis limit (GMRESmaxiterations <= MAXK - 1)
constexpr int RED_BLOCKS = 2048;

// bytes of device memory held by the library's buffers (all handles of the process)
std::atomic<long long> &device_bytes();
// bytes allocated minus bytes freed by the library's buffers on this host thread: a handle's own
// footprint is the change over its creation, whatever other threads allocate or free meanwhile
long long &thread_bytes();

// Process-wide lock held while a stream is captured into a graph and while a handle is created,
// destroyed or has state copied in or out (allocation, frees and blocking copies on other threads must
// not run during a capture).
std::mutex &capture_lock();
// Row pitch (elements) of the coarsest level's stored inverse: n rounded up to a multiple of 8.
inline int minv_pitch(int n) { return (n + 7) & ~7; }
// The capture lock as held by ow3d_create on this thread (nullptr elsewhere): set-up releases it around
// pure host work (the coarsest level's dense inverse), during which it makes no HIP call, so other
// handles can capture meanwhile.
std::unique_lock<std::mutex> *&create_lock();

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    ~DevBuf() { release(); }
    void release() {
        if (p) { (void)hipFree(p); device_bytes() -= (long long)bytes(); thread_bytes() -= (long long)bytes(); }
        p = nullptr; n = 0;
    }
    size_t bytes() const { return (n ? n : 1) * sizeof(T); }
    hipError_t alloc(size_t count) {
        release();
        n = count;
        hipError_t e = hipMalloc((void **)&p, bytes());
        if (e == hipSuccess) { device_bytes() += (long long)bytes(); thread_bytes() += (long long)bytes(); }
        else { p = nullptr; n = 0; }
        return e;
    }
    hipError_t upload(const T *h, size_t count, hipStream_t s) {
        return hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s);
    }
};

template <typename T>
struct MgLevel {
    Grid g;
    DevBuf<T> sx, sy, sz, zs, col, x, x2, b, r, minv;   // x2: second iterate buffer of out-of-place sweeps
    DevBuf<__half> minv16;                             // coarsest-level inverse in FP16 (mg_precision 2)
    int minv_pitch = 0;                                // row pitch (elements) of minv / minv16
    DevBuf<double> xf, yf, zf;   // level coordinates (restriction / prolongation weights)
    // prolongation onto this level from the next coarser one when that one is coarsened in x and y but not
    // in sigma: per fine column i (row j) the two coarse columns (rows) it interpolates from, -1 outside the
    // interior, and the weights (left, right, 1 / sum) in T (left = right index for points on a coarse line,
    // weights (1, 0, 1))
    DevBuf<int2> pxi, pyi;
    DevBuf<T> pxw, pyw;
    bool ptab = false;
    // restriction from this level onto the next coarser one in the same case: per coarse column (row) the
    // weights of the 3-point smoothing on the fine grid, (d_lo, d_lo + d_hi, d_hi, 1 / (2 (d_lo + d_hi))) in T
    DevBuf<T> rxw, ryw;
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
    DevBuf<double> sx, sy, sz, zs, zr, bath, h, cf;   // zr: sigma coordinates with the ghost level's
    MvArgs mv{};
    // surface state
    DevBuf<double> E, P, W, Ex, Exx, Ey, Eyy, Px, Py, Et, Pt, Wt, KE, KP;
    // Laplace solve
    DevBuf<double> phi, rhs, last0, last1, w, z, V, part, red, hv;
    DevBuf<double> Zf;                  // flexible GMRES: the preconditioned basis vectors
    bool flexible = false;              // right-preconditioned (flexible) GMRES instead of left
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
    // launch choices (ow3d_tune): matvec variant 0 column kernel, widest tile / 1 column kernel, 8-row
    // tiles / 2 point kernel; zebra-sweep block size 0 auto (1024 threads below 256 tiles) / 256 / 1024
    int mv_variant = 0;
    int rb_threads = 0;
    BottomArgs<double> bot64{};
    BottomArgs<float> bot32{};
    DevBuf<BottomArgs<double>> bot64d;   // device copies read by the bottom kernel
    DevBuf<BottomArgs<float>> bot32d;
    // device-driven solves: scalar state, statistics and the done flag every solve kernel reads
    DevBuf<KrylovState> ks;
    const int *done() const { return &ks.p->done; }
    const int *skip2() const { return &ks.p->skip2; }
    long tstep = 0;
    long long bytes = 0;                // device memory of this handle
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_done = nullptr;
    hipEvent_t ev_order = nullptr;      // orders a caller's stream with the handle's (device-pointer entry points)
    bool use_graph = true;
    bool capturing = false;             // a whole RK4 step is being captured (no nested graph launches)
    hipGraphExec_t pg_exec = nullptr;   // captured preconditioner (V cycles)
    const void *mg_out = nullptr;       // finest-level buffer holding the preconditioner result
    // Speculative RK4 steps: each stage's solve is enqueued with slots[stage] iteration slots (recent
    // use plus a margin) instead of its whole budget; a step in which a solve used all its slots without
    // stopping is rolled back from the snapshot (bE, bP, bW, bphi) and replayed with the whole budget.
    DevBuf<double> bE, bP, bW, bphi;
    int *report = nullptr, *report_dev = nullptr;   // pinned: overflow, amax[4] of the last step
    int slots[4] = {0, 0, 0, 0};
    std::vector<std::array<int, 4>> hist;           // slot use of recent steps
    long long replays = 0;
    bool speculate = true;
    int margin = 2;
    std::map<std::array<int, 2>, hipGraphExec_t> graphs;   // (stage, slots) -> captured stage
    double graph_dt = 0.0;
    long long captures = 0;
    double capture_ms = 0.0;
    // asynchronous stepping
    int async_left = 0;
    double async_dt = 0.0;
    bool async_busy = false;
    void drop_step_graphs() {
        for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
    }
    ~ow3d_tank() {
        drop_step_graphs();
        if (report) (void)hipHostFree(report);
        if (pg_exec) (void)hipGraphExecDestroy(pg_exec);
        if (ev_done) (void)hipEventDestroy(ev_done);
        if (ev_order) (void)hipEventDestroy(ev_order);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (s) (void)hipStreamDestroy(s);
    }
};

namespace ow3d_host {
// high-order operator (ow3d_matvec.hip); skip: the solve's done flag for launches inside a solve
int launch_matvec(const ow3d_tank *t, const double *phi, double *y, hipStream_t s, const int *skip = nullptr);
bool matvec_supported(int nzg, int A);
// multigrid (ow3d_mg.hip): hierarchy set-up for t->cfg and u = M^-1 r on device vectors
int mg_setup(ow3d_tank *t);
int precond(ow3d_tank *t, const double *r, double *u, hipStream_t s);
// per-precision parts (ow3d_mg64.hip, ow3d_mg32.hip)
int mg_build64(ow3d_tank *t);
int mg_build32(ow3d_tank *t);
int mg_fill_bottom64(ow3d_tank *t);
int mg_fill_bottom32(ow3d_tank *t);
int precond_body64(ow3d_tank *t, const double *r, const void **out, hipStream_t s, double *u);
int precond_body32(ow3d_tank *t, const double *r, const void **out, hipStream_t s, double *u);
}  // namespace ow3d_host
