// C-ABI implementation (include/ow3d.h) of the MI355X Laplace-solve-per-RK4-stage path: device
// buffers, the host-driven GMRES / defect-correction solvers and the RK4 orchestration.  Kernels live
// in ow3d_kernels.hpp, set-up geometry in ow3d_grid.hpp, the operator dispatch in ow3d_matvec.hip and
// the multigrid preconditioner in ow3d_mg.hip.
#include "ow3d_internal.hpp"

#include <chrono>
#include <thread>

std::string &ow3d_host::err() {
    static thread_local std::string e;
    return e;
}

extern "C" const char *ow3d_last_error(void) { return ow3d_host::err().c_str(); }

std::mutex &ow3d_host::capture_lock() {
    static std::mutex m;
    return m;
}

std::unique_lock<std::mutex> *&ow3d_host::create_lock() {
    static thread_local std::unique_lock<std::mutex> *l = nullptr;
    return l;
}

std::atomic<long long> &ow3d_host::device_bytes() {
    static std::atomic<long long> b{0};
    return b;
}

long long &ow3d_host::thread_bytes() {
    static thread_local long long b = 0;
    return b;
}

namespace {

// ---------------------------------------------------------------------------------------------
// Krylov solves, enqueued on the tank's stream with no host round trip (scalar work in the one-thread
// kernels of ow3d_kernels.hpp); results are read back by the entry points that need them.
// ---------------------------------------------------------------------------------------------
int red_blocks(size_t n) { return (int)std::min<size_t>(RED_BLOCKS, std::max<size_t>(1, (n + 1023) / 1024)); }

// Krylov kernel dispatch on the number of basis vectors (1 .. MAXK-1).
#define OW3D_K_CASES(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) \
    X(17) X(18) X(19) X(20) X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28) X(29) X(30) X(31)
#define OW3D_K_CASES_2(X) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) \
    X(17) X(18) X(19) X(20) X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28) X(29) X(30) X(31)
#define X2_16(X) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16)
#define X1_16(X) X(1) X2_16(X)
static_assert(MAXK == 32, "Krylov dispatch covers 1..31 vectors");

int launch_mdot(int k, bool norm, int nb, size_t n, const double *V, const double *w, double *part, const int *skip, hipStream_t s) {
    switch (k) {
#define OW3D_C(K) case K: if (norm) hipLaunchKernelGGL((k_mdot<K, 1, MAXK>), dim3(nb), dim3(256), 0, s, n, V, w, part, skip); \
                          else hipLaunchKernelGGL((k_mdot<K, 0, MAXK>), dim3(nb), dim3(256), 0, s, n, V, w, part, skip); return 0;
        OW3D_K_CASES(OW3D_C)
#undef OW3D_C
    }
    err() = "Krylov dimension out of range";
    return -1;
}
int launch_gs_update_dot(int k, int nb, size_t n, const double *V, const double *h, double *v, double *part, const int *skip,
                         hipStream_t s) {
    switch (k) {
#define OW3D_C(K) case K: hipLaunchKernelGGL((k_gs_update_dot<K, MAXK>), dim3(nb), dim3(256), 0, s, n, V, h, v, part, skip); return 0;
        OW3D_K_CASES(OW3D_C)
#undef OW3D_C
    }
    err() = "Krylov dimension out of range";
    return -1;
}
int launch_dcgs_a(int k, int nb, size_t n, double *V, const KrylovState *S, double *part, const int *skip, hipStream_t s) {
    if (n % 2 == 0 && k <= 16) {   // paired loads (16 bytes a lane): 1 % faster RK4 steps on the 513^2 workload
        switch (k) {
#define OW3D_C(K) case K: hipLaunchKernelGGL((k_dcgs_a2<K, MAXK>), dim3(nb), dim3(256), 0, s, n, V, S, part, skip); return 0;
            X2_16(OW3D_C)
#undef OW3D_C
        }
    }
    switch (k) {
#define OW3D_C(K) case K: hipLaunchKernelGGL((k_dcgs_a<K, MAXK>), dim3(nb), dim3(256), 0, s, n, V, S, part, skip); return 0;
        OW3D_K_CASES_2(OW3D_C)
#undef OW3D_C
    }
    err() = "Krylov dimension out of range";
    return -1;
}
int launch_dcgs_b(int k, int nb, size_t n, const double *V, const double *h, const KrylovState *S, double *v, double *part,
                  const int *skip, hipStream_t s) {
    if (n % 2 == 0 && k <= 16) {   // paired loads (16 bytes a lane): 1 % faster RK4 steps on the 513^2 workload
        switch (k) {
#define OW3D_C(K) case K: hipLaunchKernelGGL((k_dcgs_b2<K, MAXK>), dim3(nb), dim3(256), 0, s, n, V, h, S, v, part, skip); return 0;
            X1_16(OW3D_C)
#undef OW3D_C
        }
    }
    switch (k) {
#define OW3D_C(K) case K: hipLaunchKernelGGL((k_dcgs_b<K, MAXK>), dim3(nb), dim3(256), 0, s, n, V, h, S, v, part, skip); return 0;
        OW3D_K_CASES(OW3D_C)
#undef OW3D_C
    }
    err() = "Krylov dimension out of range";
    return -1;
}
int launch_gs_update(int k, int nb, size_t n, const double *V, const double *h, const double *scale, double *v, const int *skip,
                     hipStream_t s) {
    switch (k) {
#define OW3D_C(K) case K: hipLaunchKernelGGL((k_gs_update<K>), dim3(nb), dim3(256), 0, s, n, V, h, -1.0, 1.0, scale, v, skip); return 0;
        OW3D_K_CASES(OW3D_C)
#undef OW3D_C
    }
    err() = "Krylov dimension out of range";
    return -1;
}

// Workspace of one Laplace solve: the Krylov basis V (m + 1 vectors of n), the residual w and correction z
// vectors, the partial sums of the reductions (RED_BLOCKS x MAXK) and their totals hv (2 x MAXK: the
// two reductions of a step).
struct SolveWork {
    double *V, *w, *z, *part, *hv;
    double *Z;   // flexible GMRES: z_j = M^-1 v_j (m vectors of n)
};
SolveWork handle_work(ow3d_tank *t) { return {t->V.p, t->w.p, t->z.p, t->part.p, t->hv.p, t->Zf.p}; }

// hv[0..k-1] (with second: hv[MAXK..MAXK+k-1]) = sum over the partial sums of the last reduction
void reduce(const SolveWork &W, int k, int nb, const int *skip, hipStream_t s, bool second = false) {
    hipLaunchKernelGGL((k_reduce_parts<MAXK>), dim3(k), dim3(256), 0, s, nb, k, (const double *)W.part, W.hv + (second ? MAXK : 0), skip);
}

// Left-preconditioned GMRES(m), one cycle (iterative_solution.f90 with SPARSKIT2 gmres), orthogonalised
// by one-reduction classical Gram-Schmidt with reorthogonalisation (k_dcgs_a / k_dcgs_b / k_dcgs_column:
// every step projects twice, in two reads of the basis).  At most m operator applications (the initial
// residual and m - 1 Arnoldi steps), as the reference's ipar(5) = ipar(6) = m setting.  With t->flexible,
// right preconditioning with the preconditioned vectors kept (FGMRES, Saad 1993): the basis spans A z_j
// with z_j = M^-1 v_j, x = x0 + Z y, and the stopping test is on the true residual, whose attainable
// accuracy does not depend on how exactly M^-1 is linear; the preconditioner must then see each basis
// vector itself, so that path uses classical Gram-Schmidt with the DGKS second pass when the projection
// removed more than half of the squared norm.
int full_slots(const ow3d_tank *t) {
    if (t->cfg.solver == 0) return std::max(1, t->cfg.gmres_m);
    const int m = t->cfg.gmres_m > 1 ? t->cfg.gmres_m : 15;
    return std::min(m - 1, MAXK - 2);
}

// slots: Arnoldi steps enqueued (at most full_slots); stage: RK4 stage (0..3) the slot use is kept for.
int gmres(ow3d_tank *t, const SolveWork &W, const double *b, double *x, hipStream_t s, int slots, int stage) {
    const size_t n = t->g.n;
    const int m = t->cfg.gmres_m > 1 ? t->cfg.gmres_m : 15;
    double *V = W.V, *w = W.w;
    KrylovState *S = t->ks.p;
    const int *done = t->done(), *skip2 = t->skip2();
    const int nbl = grid_blocks(n), nb = red_blocks(n);
    const bool flex = t->flexible;
    hipLaunchKernelGGL(k_solve_begin, dim3(1), dim3(64), 0, s, S);
    if (flex) {   // v0 = b - A x0
        if (launch_matvec(t, x, V, s, done)) { err() = "matvec dispatch failed"; return -1; }
        hipLaunchKernelGGL(k_bminus, dim3(nbl), dim3(256), 0, s, n, b, V, done);
    } else {      // v0 = M^-1 (b - A x0)
        if (launch_matvec(t, x, w, s, done)) { err() = "matvec dispatch failed"; return -1; }
        hipLaunchKernelGGL(k_bminus, dim3(nbl), dim3(256), 0, s, n, b, w, done);
        if (precond(t, w, V, s)) return -1;
    }
    if (launch_mdot(1, false, nb, n, V, V, W.part, done, s)) return -1;
    reduce(W, 1, nb, done, s);
    hipLaunchKernelGGL(k_gmres_start, dim3(1), dim3(64), 0, s, S, (const double *)W.hv, t->cfg.reltol, t->cfg.abstol);
    hipLaunchKernelGGL(k_scale_state, dim3(nbl), dim3(256), 0, s, n, (const KrylovState *)S, V);
    for (int k = 1; k <= slots; k++) {
        double *vk = V + (size_t)(k - 1) * n, *vn = V + (size_t)k * n;
        if (!flex) {   // one-reduction CGS2: the operator is applied to the pending candidate in slot k - 1
            if (launch_matvec(t, vk, w, s, done)) { err() = "matvec dispatch failed"; return -1; }
            if (precond(t, w, vn, s)) return -1;
            if (k == 1) { if (launch_mdot(1, false, nb, n, V, vn, W.part, done, s)) return -1; }
            else if (launch_dcgs_a(k, nb, n, V, S, W.part, done, s)) return -1;
            reduce(W, k, nb, done, s);
            if (launch_dcgs_b(k, nb, n, V, W.hv, S, vn, W.part, done, s)) return -1;
            reduce(W, k + 1, nb, done, s, true);
            hipLaunchKernelGGL(k_dcgs_column, dim3(1), dim3(64), 0, s, S, (const double *)W.hv, (const double *)(W.hv + MAXK), k, m);
            continue;
        }
        {   // z = M^-1 v_k-1, kept; the new direction is A z
            double *zk = W.Z + (size_t)(k - 1) * n;
            if (precond(t, vk, zk, s)) return -1;
            if (launch_matvec(t, zk, vn, s, done)) { err() = "matvec dispatch failed"; return -1; }
        }
        if (launch_mdot(k, true, nb, n, V, vn, W.part, done, s)) return -1;
        reduce(W, k + 1, nb, done, s);
        hipLaunchKernelGGL(k_gmres_pass1, dim3(1), dim3(64), 0, s, S, (const double *)W.hv, k, m);
        if (launch_gs_update_dot(k, nb, n, V, W.hv, vn, W.part, skip2, s)) return -1;
        reduce(W, k + 1, nb, skip2, s);
        hipLaunchKernelGGL(k_gmres_pass2, dim3(1), dim3(64), 0, s, S, (const double *)W.hv, k, m);
        // vn = (vn - V h) / hn with the projections of the last pass
        if (launch_gs_update(k, nb, n, V, W.hv, &S->scale, vn, done, s)) return -1;
    }
    hipLaunchKernelGGL(k_gmres_finish, dim3(1), dim3(64), 0, s, S, slots, full_slots(t), stage);
    hipLaunchKernelGGL(k_basis_combine, dim3(nbl), dim3(256), 0, s, n, (const double *)(flex ? W.Z : V), (const KrylovState *)S, x);
    return 0;
}

// Defect correction (dc.f:1-151), at most m iterations (slots of them enqueued).
int dcorr(ow3d_tank *t, const SolveWork &W, const double *b, double *x, hipStream_t s, int slots, int stage) {
    const size_t n = t->g.n;
    const int nbl = grid_blocks(n), nb = red_blocks(n);
    double *w = W.w, *z = W.z;
    KrylovState *S = t->ks.p;
    const int *done = t->done();
    hipLaunchKernelGGL(k_solve_begin, dim3(1), dim3(64), 0, s, S);
    for (int it = 1; it <= slots; it++) {
        if (launch_matvec(t, x, w, s, done)) { err() = "matvec dispatch failed"; return -1; }
        hipLaunchKernelGGL(k_bminus, dim3(nbl), dim3(256), 0, s, n, b, w, done);
        if (precond(t, w, z, s)) return -1;
        if (launch_mdot(1, false, nb, n, z, z, W.part, done, s)) return -1;
        reduce(W, 1, nb, done, s);
        hipLaunchKernelGGL(k_dc_check, dim3(1), dim3(64), 0, s, S, (const double *)W.hv, it, t->cfg.reltol, t->cfg.abstol);
        hipLaunchKernelGGL(k_axpby, dim3(nbl), dim3(256), 0, s, n, 1.0, (const double *)z, 1.0, x, done);
    }
    hipLaunchKernelGGL(k_dc_finish, dim3(1), dim3(64), 0, s, S, slots, full_slots(t), stage);
    return 0;
}

// Enqueues one Laplace solve of A x = b (x: initial guess in, solution out) in workspace W (the handle's
// unless given); slots < 0: whole budget.
int solve_dev(ow3d_tank *t, const double *b, double *x, hipStream_t s, int slots = -1, int stage = 0,
              const SolveWork *W = nullptr) {
    const int f = full_slots(t);
    if (slots < 0 || slots > f) slots = f;
    const SolveWork hw = handle_work(t);
    if (!W) W = &hw;
    return t->cfg.solver == 0 ? dcorr(t, *W, b, x, s, slots, stage) : gmres(t, *W, b, x, s, slots, stage);
}

PlaneArgs plane_args(ow3d_tank *t, double *E, double *P) {
    PlaneArgs a{};
    a.nxg = t->g.nxg; a.nyg = t->g.nyg; a.nxy = t->g.nxy; a.gx = t->g.gx; a.gy = t->g.gy;
    a.a = t->g.a; a.b = t->g.b; a.nonlinear = t->cfg.nonlinear;
    a.sx = t->sx.p; a.sy = t->sy.p; a.h = t->h.p; a.bath = t->bath.p;
    a.E = E; a.P = P; a.Ex = t->Ex.p; a.Exx = t->Exx.p; a.Ey = t->Ey.p; a.Eyy = t->Eyy.p; a.Px = t->Px.p; a.Py = t->Py.p;
    a.cf = t->cf.p;
    return a;
}

void surface_dev(ow3d_tank *t, double *E, double *P, hipStream_t s) {
    PlaneArgs a = plane_args(t, E, P);
    // stencil widths of 4th, 6th and 8th order as template arguments, others at run time
    switch (2 * t->g.a + 1) {
        case 5: hipLaunchKernelGGL((k_surface_x<5>), dim3(t->g.nyg), dim3(256), 0, s, a); break;
        case 7: hipLaunchKernelGGL((k_surface_x<7>), dim3(t->g.nyg), dim3(256), 0, s, a); break;
        case 9: hipLaunchKernelGGL((k_surface_x<9>), dim3(t->g.nyg), dim3(256), 0, s, a); break;
        default: hipLaunchKernelGGL((k_surface_x<0>), dim3(t->g.nyg), dim3(256), 0, s, a); break;
    }
    if (t->g.ny > 1) {
        const dim3 grid((t->g.nxg + 63) / 64, (t->g.nyg + 3) / 4);
        switch (2 * t->g.b + 1) {
            case 5: hipLaunchKernelGGL((k_surface_y<5>), grid, dim3(256), 0, s, a); break;
            case 7: hipLaunchKernelGGL((k_surface_y<7>), grid, dim3(256), 0, s, a); break;
            case 9: hipLaunchKernelGGL((k_surface_y<9>), grid, dim3(256), 0, s, a); break;
            default: hipLaunchKernelGGL((k_surface_y<0>), grid, dim3(256), 0, s, a); break;
        }
    }
}

void wvel_dev(ow3d_tank *t, const double *phi, double *W, hipStream_t s) {
    const int rc = 2 * t->g.c + 1;
    const double *szTop = t->sz.p + (size_t)((t->g.nzg - 1) * 2) * rc;
    const dim3 grid((t->g.nxy + 255) / 256);
    switch (rc) {
        case 5: hipLaunchKernelGGL((k_vertical_velocity<5>), grid, dim3(256), 0, s, t->g.nxy, t->g.nzg, rc, szTop, (const double *)t->cf.p, phi, W); break;
        case 7: hipLaunchKernelGGL((k_vertical_velocity<7>), grid, dim3(256), 0, s, t->g.nxy, t->g.nzg, rc, szTop, (const double *)t->cf.p, phi, W); break;
        case 9: hipLaunchKernelGGL((k_vertical_velocity<9>), grid, dim3(256), 0, s, t->g.nxy, t->g.nzg, rc, szTop, (const double *)t->cf.p, phi, W); break;
        default: hipLaunchKernelGGL((k_vertical_velocity<0>), grid, dim3(256), 0, s, t->g.nxy, t->g.nzg, rc, szTop, (const double *)t->cf.p, phi, W); break;
    }
}

}  // namespace

// Extrapolated initial guesses (Runge_Kutta_4.f90:71-73, 198-200).
namespace ow3d {
__global__ void k_extrap1(size_t n, const double *__restrict__ Lnew, const double *__restrict__ Llast, double *__restrict__ phi) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        phi[q] = 1.5 * Lnew[q] - Llast[q] * 0.5;
}
__global__ void k_extrap3(size_t n, const double *__restrict__ Lnew, const double *__restrict__ Llast, double *__restrict__ phi) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        phi[q] = 2.66667 * phi[q] - 2.0 * Lnew[q] + 0.333333 * Llast[q];
}
__global__ void k_zero(size_t n, double *__restrict__ v) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) v[q] = 0.0;
}
}  // namespace ow3d

// Workspace of ow3d_solve_noalloc: a static device array, one copy per GPU the library runs on, held
// until the process ends.  Host code reaches it only through the address hipGetSymbolAddress gives for the
// current device (the array's name is a host-side stand-in, not a device address).
namespace {
constexpr size_t SCRATCH_BYTES = size_t(64) << 20;
constexpr size_t SCRATCH_DOUBLES = SCRATCH_BYTES / sizeof(double);
constexpr int SCRATCH_MAX_DEVICES = 64;
}  // namespace
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

namespace {
std::mutex scratch_mu;
double *scratch_dev[SCRATCH_MAX_DEVICES] = {};

// The scratch array's address on the current device (looked up once per device).
int scratch_ptr(double **out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) { err() = "device index beyond the scratch table"; return -1; }
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_dev[dev]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        scratch_dev[dev] = static_cast<double *>(p);
    }
    *out = scratch_dev[dev];
    return 0;
}

// Vectors in the scratch array start on 256-byte boundaries.
inline size_t pad32(size_t n) { return (n + 31) & ~size_t(31); }
}  // namespace

// ---------------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------------
extern "C" int ow3d_create(const ow3d_config *cfg, const double *h, const double *hx, const double *hy, const double *hxx,
                           const double *hyy, ow3d_tank **out) {
    if (!cfg || !h || !out) { err() = "null argument"; return -1; }
    std::unique_lock<std::mutex> lk(capture_lock());
    struct Published {   // create_lock() names lk for the duration of this call
        explicit Published(std::unique_lock<std::mutex> *l) { create_lock() = l; }
        ~Published() { create_lock() = nullptr; }
    } published(&lk);
    *out = nullptr;
    const ow3d_config &c = *cfg;
    if (c.nx < 2 || c.ny < 1 || c.nz < 3) { err() = "need Nx >= 2, Ny >= 1, Nz >= 3"; return -1; }
    if (c.ny > 1 && (c.alpha != c.beta || c.alpha != c.gamma)) { err() = "alpha, beta, gamma must be equal"; return -1; }
    if (c.ny == 1 && c.alpha != c.gamma) { err() = "alpha and gamma must be equal"; return -1; }
    if (c.gmres_m >= MAXK) { err() = "GMRESmaxiterations must be < 32"; return -1; }
    if (c.mg_precision < 0 || c.mg_precision > 2) { err() = "mg_precision must be 0 (FP64), 1 (FP32) or 2 (FP32, FP16 coarsest inverse)"; return -1; }
    if (c.precond != 0 && c.precond != 3) { err() = "precond must be 0 or 3"; return -1; }
    if (c.gmres_side < 0 || c.gmres_side > 1) { err() = "gmres_side must be 0 (left) or 1 (flexible right)"; return -1; }
    if (c.pressure_term != 0 || c.pressure_damping != 0) {
        err() = "applied free-surface pressure and pressure damping (funPressureTerm.f90, PressureTermOnOff / "
                "PDampingOnOff) are not implemented: pressure_term and pressure_damping must be 0";
        return -1;
    }
    auto t = std::make_unique<ow3d_tank>();
    t->cfg = c;
    {
        const char *ng = std::getenv("OW3D_NO_GRAPH");
        t->use_graph = !(ng && ng[0] == '1');
        const char *sp = std::getenv("OW3D_SPECULATE");   // 0: every solve enqueued with its whole budget
        t->speculate = !(sp && sp[0] == '0');
        // slots = most recently used + margin (a negative margin forces replays: used by the tests)
        if (const char *mg = std::getenv("OW3D_SLOT_MARGIN")) t->margin = std::atoi(mg);
        if (const char *mv = std::getenv("OW3D_MV")) t->mv_variant = mv[0] == 'p' ? 2 : 0;
    }
    Grid &g = t->g;
    g.nx = c.nx; g.ny = c.ny; g.nz = c.nz; g.gx = 1; g.gy = c.ny > 1 ? 1 : 0;
    g.a = c.alpha; g.b = c.ny > 1 ? c.beta : 0; g.c = c.gamma;
    g.alloc();
    if (2 * g.c + 1 > g.nzg) { err() = "sigma stencil wider than the column"; return -1; }
    if (2 * g.a + 1 > g.nx || (c.ny > 1 && 2 * g.b + 1 > g.ny)) { err() = "horizontal stencil wider than the grid"; return -1; }
    t->A = c.alpha;
    if (!matvec_supported(g.nzg, t->A)) { err() = "unsupported (Nz+1, alpha) combination for the matvec kernel"; return -1; }
    fine_coords(g, c.lx, c.ly, c.lz, c.gridz);
    const size_t nxy = g.nxy, n = g.n;
    std::copy(h, h + nxy, g.h.begin());
    if (hx) std::copy(hx, hx + nxy, g.hx.begin());
    if (hy) std::copy(hy, hy + nxy, g.hy.begin());
    if (hxx) std::copy(hxx, hxx + nxy, g.hxx.begin());
    if (hyy) std::copy(hyy, hyy + nxy, g.hyy.begin());
    grid_stencils(g);
    const long long bytes0 = thread_bytes();
    // non-blocking: set-up copies of a new handle (null stream) do not wait for other handles' work
    HIPCHK(hipStreamCreateWithFlags(&t->s, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&t->ev0));
    HIPCHK(hipEventCreate(&t->ev1));
    HIPCHK(hipEventCreateWithFlags(&t->ev_done, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&t->ev_order, hipEventDisableTiming));
    hipStream_t s = t->s;
    auto upd = [&](DevBuf<double> &d, const std::vector<double> &v) -> hipError_t {
        hipError_t e = d.alloc(v.size());
        if (e != hipSuccess) return e;
        return hipMemcpy(d.p, v.data(), v.size() * 8, hipMemcpyHostToDevice);
    };
    std::vector<double> sy = g.sy;
    if (c.ny == 1) sy.assign(6, 0.0);
    std::vector<double> zs(g.z); zs[0] = 0.0;
    HIPCHK(upd(t->sx, g.sx)); HIPCHK(upd(t->sy, sy)); HIPCHK(upd(t->sz, g.sz)); HIPCHK(upd(t->zs, zs)); HIPCHK(upd(t->h, g.h));
    HIPCHK(upd(t->zr, g.z));
    std::vector<double> bath((size_t)3 * nxy);
    for (size_t p = 0; p < nxy; p++) {
        bath[p] = g.hx[p]; bath[nxy + p] = c.ny > 1 ? g.hy[p] : 0.0; bath[2 * nxy + p] = g.hxx[p] + (c.ny > 1 ? g.hyy[p] : 0.0);
    }
    HIPCHK(upd(t->bath, bath));
    // sigma-coefficient planes of the still surface (the same expressions the surface kernels evaluate)
    std::vector<double> cf0((size_t)7 * nxy, 0.0);
    for (size_t p = 0; p < nxy; p++) {
        const double id = 1.0 / g.h[p], hx = bath[p], hy = bath[nxy + p], hs = bath[2 * nxy + p];
        const double H2 = hx * hx + hy * hy;
        cf0[p] = hx * id; cf0[nxy + p] = -hx * id; cf0[2 * nxy + p] = hy * id; cf0[3 * nxy + p] = -hy * id;
        cf0[4 * nxy + p] = (-2.0 * H2) * (id * id) + hs * id; cf0[5 * nxy + p] = (2.0 * H2) * (id * id) - hs * id;
        cf0[6 * nxy + p] = id;
    }
    HIPCHK(upd(t->cf, cf0));
    for (DevBuf<double> *b : {&t->E, &t->P, &t->W, &t->Ex, &t->Exx, &t->Ey, &t->Eyy, &t->Px, &t->Py, &t->Et, &t->Pt, &t->Wt, &t->KE, &t->KP}) {
        HIPCHK(b->alloc(nxy));
        HIPCHK(hipMemset(b->p, 0, nxy * 8));
    }
    for (DevBuf<double> *b : {&t->bE, &t->bP, &t->bW}) HIPCHK(b->alloc(nxy));
    HIPCHK(t->bphi.alloc(n));
    HIPCHK(hipHostMalloc((void **)&t->report, 8 * sizeof(int), hipHostMallocMapped | hipHostMallocCoherent));
    std::memset(t->report, 0, 8 * sizeof(int));
    HIPCHK(hipHostGetDevicePointer((void **)&t->report_dev, t->report, 0));
    for (DevBuf<double> *b : {&t->phi, &t->rhs, &t->last0, &t->last1, &t->w, &t->z, &t->io_a, &t->io_b}) {
        HIPCHK(b->alloc(n));
        HIPCHK(hipMemset(b->p, 0, n * 8));
    }
    const int m = c.gmres_m > 1 ? c.gmres_m : 15;
    HIPCHK(t->part.alloc((size_t)RED_BLOCKS * MAXK));
    HIPCHK(t->hv.alloc(2 * MAXK));
    HIPCHK(t->V.alloc((size_t)(m + 1) * n));
    HIPCHK(hipMemset(t->V.p, 0, (size_t)(m + 1) * n * 8));   // unused basis vectors stay finite
    t->flexible = c.solver == 1 && c.gmres_side == 1;
    if (t->flexible) {
        HIPCHK(t->Zf.alloc((size_t)m * n));
        HIPCHK(hipMemset(t->Zf.p, 0, (size_t)m * n * 8));
    }
    {
        KrylovState k0{};
        k0.it_min = -1;
        HIPCHK(t->ks.alloc(1));
        HIPCHK(hipMemcpy(t->ks.p, &k0, sizeof(k0), hipMemcpyHostToDevice));
    }
    t->mv.nxg = g.nxg; t->mv.nyg = g.nyg; t->mv.nxy = g.nxy;
    t->mv.sx = t->sx.p; t->mv.sy = t->sy.p; t->mv.sz = t->sz.p; t->mv.zs = t->zs.p; t->mv.cf = t->cf.p;
    t->mv.skip = nullptr;
    if (c.precond == 3 && mg_setup(t.get())) return -1;
    for (int &v : t->slots) v = full_slots(t.get());
    HIPCHK(hipStreamSynchronize(nullptr));
    HIPCHK(hipStreamSynchronize(t->s));
    t->bytes = thread_bytes() - bytes0;
    (void)s;
    *out = t.release();
    return 0;
}

extern "C" void ow3d_destroy(ow3d_tank *t) {
    if (!t) return;
    std::lock_guard<std::mutex> lk(capture_lock());
    (void)hipStreamSynchronize(t->s);
    delete t;
}

extern "C" int ow3d_dims(const ow3d_tank *t, int *nxg, int *nyg, int *nzg, int *mg_levels) {
    if (!t) return -1;
    if (nxg) *nxg = t->g.nxg;
    if (nyg) *nyg = t->g.nyg;
    if (nzg) *nzg = t->g.nzg;
    if (mg_levels) *mg_levels = t->nlev;
    return 0;
}

extern "C" void *ow3d_stream(ow3d_tank *t) { return t ? (void *)t->s : nullptr; }

extern "C" int ow3d_laplace_apply_device(ow3d_tank *t, const double *d_phi, double *d_y, void *stream) {
    if (!t) return -1;
    if (!d_phi || !d_y) { err() = "null argument"; return -1; }
    // the matvec kernels read phi's columns with 16-byte loads
    if (reinterpret_cast<uintptr_t>(d_phi) % 16 != 0 || reinterpret_cast<uintptr_t>(d_y) % 8 != 0) {
        err() = "d_phi must be 16-byte aligned and d_y 8-byte aligned";
        return -1;
    }
    // the operator's sigma coefficients are rewritten on the handle's stream by every surface update
    if (t->async_busy) { err() = "asynchronous steps still in flight"; return -1; }
    hipStream_t s = stream ? (hipStream_t)stream : t->s;
    if (s != t->s) {   // the caller's stream waits for the handle's work, and the handle's for this matvec
        HIPCHK(hipEventRecord(t->ev_order, t->s));
        HIPCHK(hipStreamWaitEvent(s, t->ev_order, 0));
    }
    if (launch_matvec(t, d_phi, d_y, s)) { err() = "matvec dispatch failed"; return -1; }
    HIPCHK(hipGetLastError());
    if (s != t->s) {
        HIPCHK(hipEventRecord(t->ev_order, s));
        HIPCHK(hipStreamWaitEvent(t->s, t->ev_order, 0));
    }
    return 0;
}

extern "C" int ow3d_laplace_apply(ow3d_tank *t, const double *phi, double *y) {
    if (!t || !phi || !y) { err() = "null argument"; return -1; }
    const size_t n = t->g.n;
    HIPCHK(hipMemcpyAsync(t->io_a.p, phi, n * 8, hipMemcpyHostToDevice, t->s));
    if (launch_matvec(t, t->io_a.p, t->io_b.p, t->s)) { err() = "matvec dispatch failed"; return -1; }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(y, t->io_b.p, n * 8, hipMemcpyDeviceToHost, t->s));
    HIPCHK(hipStreamSynchronize(t->s));
    return 0;
}

extern "C" int ow3d_time_matvec(ow3d_tank *t, int n_launches, double *avg_ms) {
    if (!t || n_launches < 1) return -1;
    HIPCHK(hipEventRecord(t->ev0, t->s));
    for (int r = 0; r < n_launches; r++)
        if (launch_matvec(t, (r & 1) ? t->io_b.p : t->io_a.p, (r & 1) ? t->io_a.p : t->io_b.p, t->s)) return -1;
    HIPCHK(hipEventRecord(t->ev1, t->s));
    HIPCHK(hipEventSynchronize(t->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, t->ev0, t->ev1));
    *avg_ms = (double)ms / n_launches;
    return 0;
}

extern "C" int ow3d_precond_apply(ow3d_tank *t, const double *r, double *u) {
    if (!t || !r || !u) { err() = "null argument"; return -1; }
    const size_t n = t->g.n;
    HIPCHK(hipMemcpyAsync(t->w.p, r, n * 8, hipMemcpyHostToDevice, t->s));
    hipLaunchKernelGGL(k_solve_begin, dim3(1), dim3(64), 0, t->s, t->ks.p);   // clears the done flag
    if (precond(t, t->w.p, t->io_b.p, t->s)) return -1;
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(u, t->io_b.p, n * 8, hipMemcpyDeviceToHost, t->s));
    HIPCHK(hipStreamSynchronize(t->s));
    return 0;
}

extern "C" int ow3d_solve(ow3d_tank *t, const double *rhs, double *sol, int *iters, double *res) {
    if (!t || !rhs || !sol) { err() = "null argument"; return -1; }
    const size_t n = t->g.n;
    HIPCHK(hipMemcpyAsync(t->io_a.p, rhs, n * 8, hipMemcpyHostToDevice, t->s));
    HIPCHK(hipMemcpyAsync(t->io_b.p, sol, n * 8, hipMemcpyHostToDevice, t->s));
    if (solve_dev(t, t->io_a.p, t->io_b.p, t->s)) return -1;
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(sol, t->io_b.p, n * 8, hipMemcpyDeviceToHost, t->s));
    KrylovState k{};
    HIPCHK(hipMemcpyAsync(&k, t->ks.p, sizeof(k), hipMemcpyDeviceToHost, t->s));
    HIPCHK(hipStreamSynchronize(t->s));
    if (iters) *iters = k.matvecs;
    if (res) *res = k.est;
    return k.status;
}

// ow3d_solve with its workspace in the static scratch array: the reduction partial sums and totals, the
// staged right-hand side and solution, the residual and correction vectors and the Krylov basis, laid out
// in that order.  The V cycle is enqueued inline (the handle's captured preconditioner graph would be
// instantiated on first use).  No device memory is allocated or freed.
extern "C" int ow3d_solve_noalloc(ow3d_tank *t, const double *rhs, double *sol, int *iters, double *res) {
    if (!t || !rhs || !sol) { err() = "null argument"; return -1; }
    if (t->async_busy) { err() = "asynchronous steps still in flight"; return -1; }
    const size_t n = t->g.n, np = pad32(n);
    const int m = t->cfg.gmres_m > 1 ? t->cfg.gmres_m : 15;
    const size_t head = pad32((size_t)RED_BLOCKS * MAXK) + pad32(2 * MAXK);
    const size_t nbasis = (size_t)(m + 1) + (t->flexible ? (size_t)m : 0);   // V, then Z
    const size_t need = head + (4 + nbasis) * np;
    if (need > SCRATCH_DOUBLES) {
        err() = "tank too large for the no-allocation solve: its workspace needs " + std::to_string(need * 8 >> 20) +
                " MiB, the static array holds " + std::to_string(SCRATCH_BYTES >> 20) + " MiB";
        return -1;
    }
    double *base = nullptr;
    if (scratch_ptr(&base)) return -1;
    double *b = base + head, *x = b + np;
    SolveWork W{x + 3 * np, x + np, x + 2 * np, base, base + pad32((size_t)RED_BLOCKS * MAXK),
                t->flexible ? x + 3 * np + (size_t)(m + 1) * np : nullptr};
    hipStream_t s = t->s;
    // the array keeps the last call's contents: the bases start from zeros, as the handle's do
    hipLaunchKernelGGL(k_zero, dim3(grid_blocks(nbasis * np)), dim3(256), 0, s, nbasis * np, W.V);
    HIPCHK(hipMemcpyAsync(b, rhs, n * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(x, sol, n * 8, hipMemcpyHostToDevice, s));
    if (solve_dev(t, b, x, s, -1, 0, &W)) return -1;
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(sol, x, n * 8, hipMemcpyDeviceToHost, s));
    KrylovState k{};
    HIPCHK(hipMemcpyAsync(&k, t->ks.p, sizeof(k), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (iters) *iters = k.matvecs;
    if (res) *res = k.est;
    return k.status;
}

extern "C" int ow3d_set_surface(ow3d_tank *t, double *E, double *P, double *Ex, double *Exx, double *Ey, double *Eyy, double *Px,
                                double *Py) {
    if (!t || !E || !P) { err() = "null argument"; return -1; }
    const size_t nxy = t->g.nxy;
    hipStream_t s = t->s;
    HIPCHK(hipMemcpyAsync(t->Et.p, E, nxy * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(t->Pt.p, P, nxy * 8, hipMemcpyHostToDevice, s));
    surface_dev(t, t->Et.p, t->Pt.p, s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(E, t->Et.p, nxy * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(P, t->Pt.p, nxy * 8, hipMemcpyDeviceToHost, s));
    double *outs[6] = {Ex, Exx, Ey, Eyy, Px, Py};
    double *devs[6] = {t->Ex.p, t->Exx.p, t->Ey.p, t->Eyy.p, t->Px.p, t->Py.p};
    for (int q = 0; q < 6; q++)
        if (outs[q]) HIPCHK(hipMemcpyAsync(outs[q], devs[q], nxy * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

extern "C" int ow3d_vertical_velocity(ow3d_tank *t, const double *phi, double *W) {
    if (!t || !phi || !W) { err() = "null argument"; return -1; }
    HIPCHK(hipMemcpyAsync(t->io_a.p, phi, t->g.n * 8, hipMemcpyHostToDevice, t->s));
    wvel_dev(t, t->io_a.p, t->Wt.p, t->s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(W, t->Wt.p, t->g.nxy * 8, hipMemcpyDeviceToHost, t->s));
    HIPCHK(hipStreamSynchronize(t->s));
    return 0;
}

namespace {
StageArgs stage_args(ow3d_tank *t, int stage, double dt, const double *Es, const double *W, double *Eo, double *Po) {
    StageArgs a{};
    a.nxg = t->g.nxg; a.nyg = t->g.nyg; a.nxy = t->g.nxy; a.dim3 = t->g.ny > 1; a.nonlinear = t->cfg.nonlinear;
    a.stage = stage; a.g = t->cfg.g; a.dt = dt; a.zero_corners = 1;
    a.E = t->E.p; a.P = t->P.p; a.Es = Es; a.W = W; a.Ex = t->Ex.p; a.Ey = t->Ey.p; a.Px = t->Px.p; a.Py = t->Py.p;
    a.KE = t->KE.p; a.KP = t->KP.p; a.Eo = Eo; a.Po = Po;
    a.zones = t->dzones.p; a.gam = t->dgam.p; a.nzones = (int)t->zones.size();
    return a;
}
}  // namespace

extern "C" int ow3d_rhs_free_surface(ow3d_tank *t, const double *E, const double *W, double *rhsE, double *rhsP) {
    if (!t || !E || !W || !rhsE || !rhsP) { err() = "null argument"; return -1; }
    const size_t nxy = t->g.nxy;
    hipStream_t s = t->s;
    // stage-1 form with dt = 0 and E = P = 0: output = 0 + 0.5*0*k is useless, so evaluate k directly:
    // use KE/KP as outputs of a stage-1 launch (KE = k1, KP = k1P).
    HIPCHK(hipMemcpyAsync(t->io_a.p, E, nxy * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(t->io_b.p, W, nxy * 8, hipMemcpyHostToDevice, s));
    StageArgs a = stage_args(t, 1, 0.0, t->io_a.p, t->io_b.p, t->io_a.p + nxy, t->io_b.p + nxy);
    a.nzones = 0;
    a.zero_corners = 0;  // Zero_Corners belongs to the RK4 driver, not to rhsFreeSurface3D
    a.E = t->io_a.p; a.P = t->io_b.p;
    hipLaunchKernelGGL(k_stage, dim3((nxy + 255) / 256), dim3(256), 0, s, a);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(rhsE, t->KE.p, nxy * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(rhsP, t->KP.p, nxy * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

extern "C" int ow3d_add_relax_zone(ow3d_tank *t, int axis, int i0, int i1, int j0, int j1, const double *gam) {
    if (!t || !gam) { err() = "null argument"; return -1; }
    if (i0 < 0 || j0 < 0 || i1 >= t->g.nxg || j1 >= t->g.nyg || i1 < i0 || j1 < j0 || (axis != 0 && axis != 1)) {
        err() = "relaxation zone outside the grid"; return -1;
    }
    std::lock_guard<std::mutex> lk(capture_lock());
    HIPCHK(hipStreamSynchronize(t->s));
    t->drop_step_graphs();   // captured steps hold the old zone arrays
    Zone z{axis, i0, i1, j0, j1, (int)t->gam.size()};
    int len = axis == 0 ? i1 - i0 + 1 : j1 - j0 + 1;
    t->gam.insert(t->gam.end(), gam, gam + len);
    t->zones.push_back(z);
    const long long b0 = thread_bytes();
    HIPCHK(t->dzones.alloc(t->zones.size()));
    HIPCHK(hipMemcpy(t->dzones.p, t->zones.data(), t->zones.size() * sizeof(Zone), hipMemcpyHostToDevice));
    HIPCHK(t->dgam.alloc(t->gam.size()));
    HIPCHK(hipMemcpy(t->dgam.p, t->gam.data(), t->gam.size() * 8, hipMemcpyHostToDevice));
    t->bytes += thread_bytes() - b0;
    return 0;
}

extern "C" int ow3d_set_state(ow3d_tank *t, const double *E, const double *P, const double *W, const double *phi) {
    if (!t || !E || !P) { err() = "null argument"; return -1; }
    std::lock_guard<std::mutex> lk(capture_lock());
    const size_t nxy = t->g.nxy, n = t->g.n;
    hipStream_t s = t->s;
    HIPCHK(hipMemcpyAsync(t->E.p, E, nxy * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(t->P.p, P, nxy * 8, hipMemcpyHostToDevice, s));
    if (W) HIPCHK(hipMemcpyAsync(t->W.p, W, nxy * 8, hipMemcpyHostToDevice, s));
    else HIPCHK(hipMemsetAsync(t->W.p, 0, nxy * 8, s));
    if (phi) HIPCHK(hipMemcpyAsync(t->phi.p, phi, n * 8, hipMemcpyHostToDevice, s));
    else HIPCHK(hipMemsetAsync(t->phi.p, 0, n * 8, s));
    HIPCHK(hipMemsetAsync(t->last0.p, 0, n * 8, s));
    HIPCHK(hipMemsetAsync(t->last1.p, 0, n * 8, s));
    surface_dev(t, t->E.p, t->P.p, s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    t->tstep = 0;
    return 0;
}

extern "C" int ow3d_get_state(ow3d_tank *t, double *E, double *P, double *W, double *phi) {
    if (!t) return -1;
    std::lock_guard<std::mutex> lk(capture_lock());
    const size_t nxy = t->g.nxy, n = t->g.n;
    hipStream_t s = t->s;
    if (E) HIPCHK(hipMemcpyAsync(E, t->E.p, nxy * 8, hipMemcpyDeviceToHost, s));
    if (P) HIPCHK(hipMemcpyAsync(P, t->P.p, nxy * 8, hipMemcpyDeviceToHost, s));
    if (W) HIPCHK(hipMemcpyAsync(W, t->W.p, nxy * 8, hipMemcpyDeviceToHost, s));
    if (phi) HIPCHK(hipMemcpyAsync(phi, t->phi.p, n * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

static void copy_dev(const double *src, double *dst, size_t n, hipStream_t s) {
    hipLaunchKernelGGL((k_convert<double, double>), dim3(grid_blocks(n)), dim3(256), 0, s, n, src, dst, nullptr);
}

// Enqueues stage `stage` (1..4) of RK4 step number t->tstep (Runge_Kutta_4.f90:20-299): the surface
// update, the Laplace solve (`slots` iteration slots) and the vertical velocity.  Stage 1 starts with the
// step's snapshot (the state a replay starts from), stage 4 ends with the step's report to host memory.
static int stage_body(ow3d_tank *t, double dt, int stage, int slots) {
    hipStream_t s = t->s;
    const size_t nxy = t->g.nxy, n = t->g.n;
    const int nb2 = (int)((nxy + 255) / 256), nbl = grid_blocks(n);
    // LASTPHI history (Runge_Kutta_4.f90:118-121): Lnew = phi at the start of this step, Llast = phi at
    // the start of the previous one
    double *Lnew = t->last0.p, *Llast = t->last1.p;
    if (stage == 1) {
        copy_dev(t->E.p, t->bE.p, nxy, s); copy_dev(t->P.p, t->bP.p, nxy, s); copy_dev(t->W.p, t->bW.p, nxy, s);
        copy_dev(t->phi.p, t->bphi.p, n, s);
        hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(64), 0, s, t->ks.p);
        copy_dev(Lnew, Llast, n, s);
        copy_dev(t->phi.p, Lnew, n, s);
    }
    const double *Es = stage == 1 ? t->E.p : t->Et.p;
    const double *Wst = stage == 1 ? t->W.p : t->Wt.p;
    double *Eo = stage == 4 ? t->E.p : t->Et.p;
    double *Po = stage == 4 ? t->P.p : t->Pt.p;
    StageArgs a = stage_args(t, stage, dt, Es, Wst, Eo, Po);
    hipLaunchKernelGGL(k_stage, dim3(nb2), dim3(256), 0, s, a);
    surface_dev(t, Eo, Po, s);
    hipLaunchKernelGGL(k_fill_rhs, dim3(grid_blocks(n)), dim3(256), 0, s, t->g.nxg, t->g.nyg, t->g.nzg, t->g.gx, t->g.gy,
                       (const double *)Po, t->rhs.p);
    if (stage == 1 && t->cfg.extrapolation == 1)
        hipLaunchKernelGGL(k_extrap1, dim3(nbl), dim3(256), 0, s, n, (const double *)Lnew, (const double *)Llast, t->phi.p);
    if (stage == 3 && t->cfg.extrapolation == 1 && t->tstep > 1)
        hipLaunchKernelGGL(k_extrap3, dim3(nbl), dim3(256), 0, s, n, (const double *)Lnew, (const double *)Llast, t->phi.p);
    if (solve_dev(t, t->rhs.p, t->phi.p, s, slots, stage - 1)) return -1;
    wvel_dev(t, t->phi.p, stage == 4 ? t->W.p : t->Wt.p, s);
    if (stage == 4) hipLaunchKernelGGL(k_step_report, dim3(1), dim3(64), 0, s, (const KrylovState *)t->ks.p, t->report_dev);
    return 0;
}

// Enqueues step t->tstep stage by stage.  From the second step after a state change on, a stage's launch
// sequence depends only on dt, the stage and its slots: it is captured once into a hipGraph per
// (stage, slots) and replayed.
static int step_launch(ow3d_tank *t, double dt, const int *slots) {
    for (int stage = 1; stage <= 4; stage++) {
        if (!t->use_graph || t->tstep < 2) {
            if (stage_body(t, dt, stage, slots[stage - 1])) return -1;
            continue;
        }
        if (dt != t->graph_dt) { t->drop_step_graphs(); t->graph_dt = dt; }
        const std::array<int, 2> key{stage, slots[stage - 1]};
        auto it = t->graphs.find(key);
        if (it == t->graphs.end()) {
            const auto c0 = std::chrono::steady_clock::now();
            std::lock_guard<std::mutex> lk(capture_lock());
            hipGraph_t graph = nullptr;
            HIPCHK(hipStreamBeginCapture(t->s, hipStreamCaptureModeThreadLocal));
            t->capturing = true;
            const int rc = stage_body(t, dt, stage, slots[stage - 1]);
            t->capturing = false;
            hipError_t e = hipStreamEndCapture(t->s, &graph);
            if (rc) { if (graph) (void)hipGraphDestroy(graph); return -1; }
            HIPCHK(e);
            hipGraphExec_t exec = nullptr;
            e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            HIPCHK(e);
            it = t->graphs.emplace(key, exec).first;
            t->captures++;
            t->capture_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - c0).count();
        }
        HIPCHK(hipGraphLaunch(it->second, t->s));
    }
    HIPCHK(hipGetLastError());
    return 0;
}

// Next step: t->tstep advances, enqueued with the current slots.
static int step_next(ow3d_tank *t, double dt) {
    t->tstep++;
    return step_launch(t, dt, t->slots);
}

// Called once the step in flight has completed.  Rolls it back and re-enqueues it with every solve's
// whole budget when a solve ran out of slots (*replayed = true; call again when that completes);
// otherwise sets the slots of the next step from recent use.
static int step_settle(ow3d_tank *t, double dt, bool *replayed) {
    *replayed = false;
    const volatile int *r = t->report;
    const int F = full_slots(t);
    if (r[0]) {
        hipStream_t s = t->s;
        const size_t nxy = t->g.nxy, n = t->g.n;
        t->replays++;
        copy_dev(t->bE.p, t->E.p, nxy, s); copy_dev(t->bP.p, t->P.p, nxy, s); copy_dev(t->bW.p, t->W.p, nxy, s);
        copy_dev(t->bphi.p, t->phi.p, n, s);
        copy_dev(t->last1.p, t->last0.p, n, s);   // stage 1 shifts the history again
        hipLaunchKernelGGL(k_step_restore, dim3(1), dim3(64), 0, s, t->ks.p);
        surface_dev(t, t->E.p, t->P.p, s);   // plane derivatives and sigma coefficients of the restored surface
        const int full[4] = {F, F, F, F};
        for (int &v : t->slots) v = F;
        *replayed = true;
        return step_launch(t, dt, full);
    }
    if (!t->speculate) return 0;
    t->hist.push_back({r[1], r[2], r[3], r[4]});
    if (t->hist.size() > 8) t->hist.erase(t->hist.begin());
    // grow at once, shrink only by more than one slot (fewer distinct stage graphs)
    for (int q = 0; q < 4; q++) {
        int mx = 0;
        for (const auto &h : t->hist) mx = std::max(mx, h[q]);
        const int need = std::min(F, std::max(1, mx + t->margin));
        if (need > t->slots[q] || need + 1 < t->slots[q]) t->slots[q] = need;
    }
    return 0;
}

// Steps enqueued by rk4_steps_enqueue: zero the non-convergence flag of the batch of steps first.
static int status_reset(ow3d_tank *t) {
    HIPCHK(hipMemsetAsync(&t->ks.p->status_any, 0, sizeof(int), t->s));
    return 0;
}

static int steps_status(ow3d_tank *t) {
    int st = 0;
    HIPCHK(hipMemcpy(&st, &t->ks.p->status_any, sizeof(int), hipMemcpyDeviceToHost));
    return st ? 1 : 0;
}

// Waits for the step in flight and settles it (replaying as often as needed).
static int step_finish(ow3d_tank *t, double dt) {
    bool again = true;
    while (again) {
        HIPCHK(hipStreamSynchronize(t->s));
        if (step_settle(t, dt, &again)) return -1;
    }
    return 0;
}

extern "C" int ow3d_rk4_steps(ow3d_tank *t, int nsteps, double time0, double dt) {
    if (!t) return -1;
    (void)time0;
    if (t->async_busy) { err() = "asynchronous steps still in flight"; return -1; }
    if (status_reset(t)) return -1;
    for (int q = 0; q < nsteps; q++)
        if (step_next(t, dt) || step_finish(t, dt)) return -1;
    HIPCHK(hipStreamSynchronize(t->s));
    return steps_status(t);
}

extern "C" int ow3d_rk4_steps_async(ow3d_tank *t, int nsteps, double dt) {
    if (!t || nsteps < 0) { err() = "null handle or negative step count"; return -1; }
    if (t->async_busy) { err() = "asynchronous steps still in flight"; return -1; }
    if (status_reset(t)) return -1;
    t->async_left = nsteps;
    t->async_dt = dt;
    if (nsteps == 0) return 0;
    if (step_next(t, dt)) return -1;
    HIPCHK(hipEventRecord(t->ev_done, t->s));
    t->async_busy = true;
    return 0;
}

// Advances the handle's asynchronous steps as far as completed work allows: 1 when all are done.
extern "C" int ow3d_query(ow3d_tank *t) {
    if (!t) return -1;
    if (!t->async_busy) return 1;
    hipError_t e = hipEventQuery(t->ev_done);
    if (e == hipErrorNotReady) return 0;
    HIPCHK(e);
    bool replayed = false;
    if (step_settle(t, t->async_dt, &replayed)) return -1;
    if (!replayed && --t->async_left == 0) { t->async_busy = false; return 1; }
    if (!replayed && step_next(t, t->async_dt)) return -1;
    HIPCHK(hipEventRecord(t->ev_done, t->s));
    return 0;
}

extern "C" int ow3d_wait(ow3d_tank *t) {
    if (!t) return -1;
    while (true) {
        if (t->async_busy) HIPCHK(hipEventSynchronize(t->ev_done));
        const int r = ow3d_query(t);
        if (r < 0) return r;
        if (r == 1) break;
    }
    return steps_status(t);
}

extern "C" int ow3d_steps_status(ow3d_tank *t) {
    if (!t) { err() = "null handle"; return -1; }
    if (t->async_busy) { err() = "asynchronous steps still in flight"; return -1; }
    HIPCHK(hipStreamSynchronize(t->s));
    return steps_status(t);
}

extern "C" int ow3d_batch_rk4_steps(ow3d_tank *const *tanks, int ntanks, int nsteps, const double *dt) {
    if (!tanks || !dt || ntanks < 0 || nsteps < 0) { err() = "null argument"; return -1; }
    for (int i = 0; i < ntanks; i++) {
        if (!tanks[i]) { err() = "null handle in batch"; return -1; }
        if (tanks[i]->async_busy) { err() = "asynchronous steps still in flight"; return -1; }
    }
    // every tank advances on its own stream as soon as its previous step has settled (no barrier
    // between tanks); the host polls the tanks' step events
    for (int i = 0; i < ntanks; i++)
        if (ow3d_rk4_steps_async(tanks[i], nsteps, dt[i])) return -1;
    for (int left = ntanks; left > 0;) {
        left = 0;
        bool progressed = false;
        for (int i = 0; i < ntanks; i++) {
            if (!tanks[i]->async_busy) continue;
            const int before = tanks[i]->async_left;
            const int r = ow3d_query(tanks[i]);
            if (r < 0) return r;
            if (r == 0) left++;
            progressed |= (r == 1) || tanks[i]->async_left != before;
        }
        if (left > 0 && !progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    int st = 0;
    for (int i = 0; i < ntanks; i++) {
        const int r = steps_status(tanks[i]);
        if (r < 0) return r;
        st |= r;
    }
    return st;
}

extern "C" int ow3d_step_stats(const ow3d_tank *t, long long *replays, long long *captures, double *capture_ms) {
    if (!t) return -1;
    if (replays) *replays = t->replays;
    if (captures) *captures = t->captures;
    if (capture_ms) *capture_ms = t->capture_ms;
    return 0;
}

extern "C" int ow3d_stats(const ow3d_tank *t, long long *total, long long *solves, int *itmin, int *itmax) {
    if (!t) return -1;
    KrylovState k{};
    HIPCHK(hipStreamSynchronize(t->s));
    HIPCHK(hipMemcpy(&k, t->ks.p, sizeof(k), hipMemcpyDeviceToHost));
    if (total) *total = k.it_total;
    if (solves) *solves = k.it_solves;
    if (itmin) *itmin = k.it_min;
    if (itmax) *itmax = k.it_max;
    return 0;
}

extern "C" long long ow3d_device_bytes(const ow3d_tank *t) { return t ? t->bytes : -1; }

extern "C" int ow3d_kinematics(ow3d_tank *t, const double *phi, double *U, double *V, double *W) {
    if (!t || !U || !V || !W) { err() = "null argument"; return -1; }
    if (t->async_busy) { err() = "asynchronous steps still in flight"; return -1; }
    std::lock_guard<std::mutex> lk(capture_lock());
    hipStream_t s = t->s;
    const size_t n = t->g.n;
    const double *src = t->phi.p;
    if (phi) {
        HIPCHK(hipMemcpyAsync(t->io_a.p, phi, n * 8, hipMemcpyHostToDevice, s));
        src = t->io_a.p;
    }
    KinArgs a{};
    a.nxg = t->g.nxg; a.nyg = t->g.nyg; a.nzg = t->g.nzg; a.a = t->g.a; a.b = t->g.b; a.c = t->g.c;
    a.nonlinear = t->cfg.nonlinear;
    a.sx = t->sx.p; a.sy = t->sy.p; a.sz = t->sz.p; a.z = t->zr.p; a.h = t->h.p; a.bath = t->bath.p;
    a.E = t->E.p; a.Ex = t->Ex.p; a.Ey = t->Ey.p;
    // outputs: U in io_b, V in w, W in z (scratch vectors of the handle, free outside a solve)
    hipLaunchKernelGGL(k_kinematics, dim3(grid_blocks(n)), dim3(256), 0, s, a, src, t->io_b.p, t->w.p, t->z.p, (double *)nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(U, t->io_b.p, n * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(V, t->w.p, n * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(W, t->z.p, n * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

extern "C" int ow3d_kinematics_gradients(ow3d_tank *t, const double *phi, double *G) {
    if (!t || !G) { err() = "null argument"; return -1; }
    if (t->async_busy) { err() = "asynchronous steps still in flight"; return -1; }
    const size_t n = t->g.n;
    if ((size_t)t->cfg.gmres_m + 1 < 13) { err() = "the kinematics gradients need GMRESmaxiterations >= 12 (they use the Krylov basis as workspace)"; return -1; }
    std::lock_guard<std::mutex> lk(capture_lock());
    hipStream_t s = t->s;
    const double *src = t->phi.p;
    if (phi) {
        HIPCHK(hipMemcpyAsync(t->io_a.p, phi, n * 8, hipMemcpyHostToDevice, s));
        src = t->io_a.p;
    }
    KinArgs a{};
    a.nxg = t->g.nxg; a.nyg = t->g.nyg; a.nzg = t->g.nzg; a.a = t->g.a; a.b = t->g.b; a.c = t->g.c;
    a.nonlinear = t->cfg.nonlinear;
    a.sx = t->sx.p; a.sy = t->sy.p; a.sz = t->sz.p; a.z = t->zr.p; a.h = t->h.p; a.bath = t->bath.p;
    a.E = t->E.p; a.Ex = t->Ex.p; a.Ey = t->Ey.p;
    // workspace: the Krylov basis (free outside a solve) holds U, V, w, W and the 9 gradient fields
    double *U = t->V.p, *V = U + n, *Wv = V + n, *Ws = Wv + n, *Gd = Ws + n;
    hipLaunchKernelGGL(k_kinematics, dim3(grid_blocks(n)), dim3(256), 0, s, a, src, U, V, Wv, Ws);
    hipLaunchKernelGGL(k_kinematics_grad, dim3(grid_blocks(n)), dim3(256), 0, s, a, (const double *)Ws, (const double *)U,
                       (const double *)V, Gd);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(G, Gd, 9 * n * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

extern "C" int ow3d_surface_slopes(ow3d_tank *t, double *Ex, double *Ey) {
    if (!t || !Ex || !Ey) { err() = "null argument"; return -1; }
    if (t->async_busy) { err() = "asynchronous steps still in flight"; return -1; }
    const size_t nxy = t->g.nxy;
    HIPCHK(hipMemcpyAsync(Ex, t->Ex.p, nxy * 8, hipMemcpyDeviceToHost, t->s));
    HIPCHK(hipMemcpyAsync(Ey, t->Ey.p, nxy * 8, hipMemcpyDeviceToHost, t->s));
    HIPCHK(hipStreamSynchronize(t->s));
    return 0;
}

extern "C" int ow3d_tune(ow3d_tank *t, int matvec_variant, int sweep_threads) {
    if (!t) { err() = "null handle"; return -1; }
    if (matvec_variant < 0 || matvec_variant > 3) { err() = "matvec variant must be 0, 1, 2 or 3"; return -1; }
    if (sweep_threads != 0 && sweep_threads != 256 && sweep_threads != 1024) { err() = "sweep threads must be 0, 256 or 1024"; return -1; }
    if (t->async_busy) { err() = "asynchronous steps still in flight"; return -1; }
    HIPCHK(hipStreamSynchronize(t->s));
    std::lock_guard<std::mutex> lk(capture_lock());
    t->mv_variant = matvec_variant;
    t->rb_threads = sweep_threads;
    t->drop_step_graphs();   // captured with the old launch shapes
    if (t->pg_exec) { HIPCHK(hipGraphExecDestroy(t->pg_exec)); t->pg_exec = nullptr; }
    return 0;
}

// Multigrid hierarchy ow3d_create would build for cfg (host only): level count and coarsest unknowns.
extern "C" int ow3d_mg_plan(const ow3d_config *cfg, int *nlev, int *coarsest) {
    if (!cfg || cfg->nx < 2 || cfg->ny < 1 || cfg->nz < 3) { err() = "invalid configuration"; return -1; }
    Grid g;
    g.nx = cfg->nx; g.ny = cfg->ny; g.nz = cfg->nz; g.gx = 1; g.gy = cfg->ny > 1 ? 1 : 0;
    g.a = cfg->alpha; g.b = cfg->ny > 1 ? cfg->beta : 0; g.c = cfg->gamma;
    g.alloc();
    std::vector<Grid> lev;
    if (!mg_levels(g, cfg->mg_maxgrids, lev, err())) return -1;
    if (nlev) *nlev = (int)lev.size();
    if (coarsest) *coarsest = (int)lev[0].n;
    return 0;
}

// Device memory ow3d_create would allocate for cfg (the buffers it allocates, without the allocator's
// rounding): fine-grid vectors and Krylov basis, surface planes, operator tables, multigrid levels.
extern "C" long long ow3d_estimate_bytes(const ow3d_config *cfg) {
    if (!cfg || cfg->nx < 2 || cfg->ny < 1 || cfg->nz < 3) { err() = "invalid configuration"; return -1; }
    Grid g;
    g.nx = cfg->nx; g.ny = cfg->ny; g.nz = cfg->nz; g.gx = 1; g.gy = cfg->ny > 1 ? 1 : 0;
    g.a = cfg->alpha; g.b = cfg->ny > 1 ? cfg->beta : 0; g.c = cfg->gamma;
    g.alloc();
    const long long n = (long long)g.n, nxy = g.nxy;
    const int m = cfg->gmres_m > 1 ? cfg->gmres_m : 15;
    long long b = 8 * n * (8 + m + 1 + 1) + 8 * nxy * (25 + 3);   // + step snapshot (phi, E, P, W)
    const bool flex = cfg->solver == 1 && cfg->gmres_side == 1;
    if (flex) b += 8 * n * m;                                        // preconditioned basis of flexible GMRES
    b += 8LL * ((long long)g.nxg * 2 * g.ra + (cfg->ny > 1 ? (long long)g.nyg * 2 * g.rb : 6) + 2LL * g.nzg * 2 * g.rc);
    b += 8LL * RED_BLOCKS * MAXK + 16LL * MAXK + (long long)sizeof(KrylovState);
    if (cfg->precond == 3) {
        std::vector<Grid> lev;
        if (!mg_levels(g, cfg->mg_maxgrids, lev, err())) return -1;
        const long long p = cfg->mg_precision == 0 ? 8 : 4, pc = cfg->mg_precision == 2 ? 2 : p;
        for (size_t k = 0; k < lev.size(); k++) {
            const Grid &L = lev[k];
            b += p * ((long long)L.n * 4 + 6LL * L.nxy + 12LL * (L.nxg + L.nyg + L.nzg)) + 8LL * (L.nxg + L.nyg + L.nzg);
            if (k == 0) b += pc * (long long)L.n * (long long)minv_pitch((int)L.n);
        }
    }
    return b;
}
