// Learned Davidson correction: placeholder until the BF16 MLP kernels land.
#include "mlp.h"

namespace pci {

void mlp_begin(Ctx &, MlpPrecond &, int) {
    throw Error(PCI_E_UNSUPPORTED, "learned preconditioner: no trained network in this context");
}
const float *mlp_apply(Ctx &, MlpPrecond &, const double *, const double *, const double *, int, int) {
    return nullptr;
}
bool mlp_accept(Ctx &, MlpPrecond &, const std::vector<double> &, const std::vector<double> &,
                const std::vector<int> &) {
    return false;
}
void mlp_end(Ctx &, MlpPrecond &) {}

}  // namespace pci
