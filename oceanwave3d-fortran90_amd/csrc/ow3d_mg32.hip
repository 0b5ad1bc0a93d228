// Multigrid preconditioner in FP32 (level operators, smoother, transfers, bottom solve).
#include "ow3d_mg.hpp"

namespace ow3d_host {

int mg_build32(ow3d_tank *t) { return mg::mg_build<float>(t, t->mg32); }
int mg_fill_bottom32(ow3d_tank *t) { return mg::fill_bottom<float>(t, t->mg32, t->bot32, t->bot32d); }
int precond_body32(ow3d_tank *t, const double *r, const void **out, hipStream_t s, double *u) {
    return mg::precond_body<float>(t, t->mg32, r, out, s, u);
}

}  // namespace ow3d_host
