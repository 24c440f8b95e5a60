// Learned Davidson correction (MLP preconditioner) -- interface used by the
// Davidson driver.  See mlp.hip.
#pragma once
#include <vector>

#include "internal.h"

namespace pci {

struct MlpPrecond {
    bool active = false;      // learned correction in use for this iteration
    bool disabled = false;    // fell back to the diagonal for the rest of the solve
    int ldo = 0;              // leading dimension of the output block
    int nlv = 0;
    double prev_cnx = 0.0;    // residual before the last learned iteration
    DBuf<float> out;          // nd x ldo learned correction
    DBuf<float> feat;         // per-row scratch
};

// Prepare for a solve of nlv roots (the context must hold trained weights).
void mlp_begin(Ctx &c, MlpPrecond &m, int nlv);
// Learned correction for rows >= nd0 from the residual block R (nd x nlv,
// leading dimension ldr), the shifted diagonal D and the Ritz values theta.
// Returns a device pointer to the nd x m.ldo FP32 block, or nullptr when the
// diagonal correction is to be used this iteration.
const float *mlp_apply(Ctx &c, MlpPrecond &m, const double *D, const double *theta, const double *R, int ldr,
                       int nd0);
// Record the outcome of the iteration (residual norms per root); returns
// whether the learned correction was used.
bool mlp_accept(Ctx &c, MlpPrecond &m, const std::vector<double> &gram, const std::vector<double> &resid,
                const std::vector<int> &conv);
void mlp_end(Ctx &c, MlpPrecond &m);

}  // namespace pci
