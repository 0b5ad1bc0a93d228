// Multigrid preconditioner in FP64 (level operators, smoother, transfers, bottom solve).
#include "ow3d_mg.hpp"

namespace ow3d_host {

int mg_build64(ow3d_tank *t) { return mg::mg_build<double>(t, t->mg64); }
int mg_fill_bottom64(ow3d_tank *t) { return mg::fill_bottom<double>(t, t->mg64, t->bot64, t->bot64d); }
int precond_body64(ow3d_tank *t, const double *r, const void **out, hipStream_t s, double *u) {
    return mg::precond_body<double>(t, t->mg64, r, out, s, u);
}

}  // namespace ow3d_host
