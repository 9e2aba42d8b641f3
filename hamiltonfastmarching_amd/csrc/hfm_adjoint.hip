// hfm_adjoint.hip -- backward differentiation (placeholder until the adjoint kernels land).
#include "hfm_internal.h"

int hfm_plan_adjoint_impl(hfm_plan *, int, const double *, const double *, double *, int64_t *, double *, int, int *,
                          void *) {
    hfm::set_error("hfm_plan_adjoint: not available in this build");
    return 1;
}
