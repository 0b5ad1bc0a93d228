// Geometric multigrid preconditioner (multigrid/*.f90 of the reference, with a zebra z-line smoother):
// hierarchy set-up and the preconditioner entry with its hipGraph replay.  The per-precision work is in
// ow3d_mg64.hip / ow3d_mg32.hip.
#include "ow3d_internal.hpp"

namespace ow3d_host {

// Multigrid hierarchy of the tank (MGPreProcess.f90:6-460 with coarsening strategy 0), level operators in
// the precision the configuration asks for, and the levels handed to the one-workgroup bottom kernel.
int mg_setup(ow3d_tank *t) {
    const ow3d_config &c = t->cfg;
    if (!mg_levels(t->g, c.mg_maxgrids, t->levels, err())) return -1;
    t->nlev = (int)t->levels.size();
    for (auto &L : t->levels)
        if (L.nzg < 3 || L.nzg > 14) { err() = "multigrid level depth outside 3..14"; return -1; }
    if (t->levels[0].n > 4096) { err() = "coarsest multigrid level too large for the dense solve"; return -1; }
    int rc = c.mg_precision == 0 ? mg_build64(t) : mg_build32(t);
    if (rc) return -1;
    // levels small enough for the one-workgroup bottom kernel (OW3D_MG_BOTTOM: point limit, 0 = off)
    // The group must share one column depth on levels 1 .. bot_top (level 0 is solved densely).
    long lim = 0;
    if (const char *e = std::getenv("OW3D_MG_BOTTOM")) lim = std::atol(e);
    if (c.mg_precision == 2) lim = 0;   // the bottom kernel reads the FP32 inverse
    if (const char *e = std::getenv("OW3D_MG_FUSED_RB")) t->fused_rb = e[0] != '0';
    t->bot_top = -1;
    for (int k = 1; k < t->nlev && k < MG_BOTTOM_MAX; k++) {
        if ((long)t->levels[k].n > lim) break;
        bool same = true;
        for (int q = 1; q < k; q++) same &= t->levels[q].nzg == t->levels[k].nzg;
        if (same) t->bot_top = k;
    }
    return c.mg_precision == 0 ? mg_fill_bottom64(t) : mg_fill_bottom32(t);
}

// u = M^-1 r (device vectors, FP64 in / out), skipped once the solve's done flag is set.  When r is the
// handle's residual buffer the V cycle is replayed from a hipGraph captured on first use (one graph per
// handle; fixed buffers), unless a whole RK4 step is being captured (its graph then holds the V cycle).
int precond(ow3d_tank *t, const double *r, double *u, hipStream_t s) {
    const size_t n = t->g.n;
    if (t->cfg.precond != 3) {
        hipLaunchKernelGGL((k_convert<double, double>), dim3(grid_blocks(n)), dim3(256), 0, s, n, r, u, t->done());
        return 0;
    }
    if (t->use_graph && !t->capturing && r == t->w.p && s == t->s) {
        if (!t->pg_exec) {
            std::lock_guard<std::mutex> lk(capture_lock());
            hipGraph_t graph = nullptr;
            HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            int rc = (t->cfg.mg_precision == 0 ? precond_body64(t, r, &t->mg_out, s, nullptr)
                                               : precond_body32(t, r, &t->mg_out, s, nullptr));
            hipError_t e = hipStreamEndCapture(s, &graph);
            if (rc) { if (graph) (void)hipGraphDestroy(graph); return -1; }
            HIPCHK(e);
            HIPCHK(hipGraphInstantiate(&t->pg_exec, graph, nullptr, nullptr, 0));
            (void)hipGraphDestroy(graph);
        }
        HIPCHK(hipGraphLaunch(t->pg_exec, s));
    } else {
        // inline (eager, or inside a captured step): the last sweep writes u itself
        const void *o = nullptr;
        if ((t->cfg.mg_precision == 0 ? precond_body64(t, r, &o, s, u) : precond_body32(t, r, &o, s, u))) return -1;
        if (!o) return 0;
        t->mg_out = o;
    }
    if (t->cfg.mg_precision == 0)
        hipLaunchKernelGGL((k_convert<double, double>), dim3(grid_blocks(n)), dim3(256), 0, s, n, (const double *)t->mg_out, u, t->done());
    else
        hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_blocks(n)), dim3(256), 0, s, n, (const float *)t->mg_out, u, t->done());
    return 0;
}


}  // namespace ow3d_host
