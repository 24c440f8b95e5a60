// Learned Davidson correction (MLP preconditioner): network layout and the
// hooks the Davidson driver calls.  Kernels in mlp.hip.
//
// Per (determinant i >= Nd0, root j) the network maps features of the
// residual component to a gain g, and the correction is t_ij = s_ij * g with
// s = r + c, where c_ij = sum_{k < Nd0} H_ik t_kj is the coupling of the row
// to the head-block correction t (zero for pconf's plain correction).  The
// plain Dvdsn correction is g = 1/(theta_j - D_i) with c = 0
// (src/davidson.f90:482-488); the network learns a residual on top of a
// clipped version of it:  g = -f1 + MLP(x).
//
//   x[0]  u   = D_i - theta_j
//   x[1]  f1  = 1 / (u + sign(u) * 0.02)
//   x[2]  tanh(u / 0.1)
//   x[3]  log(|u| + 1e-3)
//   x[4]  excitation rank of the determinant's configuration (vs the first), / 3
//   x[5]  r_ij / rms(r_j)
//   x[6]  sign(u)
//   x[7]  c_ij / rms(r_j)
//   x[8:24]  embedding row E[bucket_i], bucket = 64 * config / Nc
//   x[24:32] 0
//   h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), y = w3 . h2 + b3
// W1 (128x32), W2 (128x128) and E (64x16) are used in BF16 with FP32
// accumulation (MFMA 32x32x8 bf16); master copies are FP32.
#pragma once
#include <vector>

#include "internal.h"

namespace pci {

constexpr int MLP_NF = 8;      // base features
constexpr int MLP_IN = 32;     // padded input width
constexpr int MLP_H = 128;     // hidden width
constexpr int MLP_NB = 64;     // embedding buckets
constexpr int MLP_NE = 16;     // embedding width
constexpr int MLP_WARMUP = 3;  // diagonal iterations before the first learned one
// flat parameter layout (FP32)
constexpr int64_t MLP_OFF_W1 = 0;
constexpr int64_t MLP_OFF_B1 = MLP_OFF_W1 + MLP_H * MLP_IN;
constexpr int64_t MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;
constexpr int64_t MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr int64_t MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;
constexpr int64_t MLP_OFF_B3 = MLP_OFF_W3 + MLP_H;
constexpr int64_t MLP_OFF_E = MLP_OFF_B3 + 1;
constexpr int64_t MLP_NPARAM = MLP_OFF_E + MLP_NB * MLP_NE;

struct MlpNet {
    bool ready = false;
    DBuf<float> w;        // FP32 master parameters
    DBuf<uint16_t> wb;    // BF16 copy of W1 | W2 | E
    DBuf<float> g;        // gradient accumulator
    DBuf<float> m, v;     // Adam moments
    int64_t step = 0;
};

struct MlpPrecond {
    bool active = false;      // learned correction in use for this iteration
    bool disabled = false;    // fell back to the diagonal for the rest of the solve
    bool tried = false;       // the learned correction was used in the last iteration
    int ldo = 0;              // leading dimension of the output block
    int nlv = 0;
    int seen = 0;             // iterations observed
    double prev_cnx = -1.0;   // residual norm of the previous iteration
    double rate_diag = -1.0;  // last residual reduction factor of a diagonal step
    DBuf<float> out;          // nd x ldo learned correction
    DBuf<float> rms;          // per-root residual rms
};

// Collection of (features, residual, target) samples during a Davidson run
// against reference eigenvectors (training data for the network).
struct MlpCollect {
    const double *xref = nullptr;   // device, nd x nlv row-major converged vectors
    int64_t cap = 0, n = 0;
    int stride = 1;                 // every stride-th row >= Nd0 is sampled
    DBuf<float> x;                  // cap x MLP_NF
    DBuf<int> bucket;               // cap
    DBuf<float> r, e;               // residual component r (+ c when coupled), aligned error component
    std::vector<float> scale;       // per sample block: c = <t_diag, e>/<e, e>
    std::vector<int64_t> block_end;
};

void mlp_begin(Ctx &c, MlpPrecond &m, int nlv);
// Learned correction for rows >= nd0 from the residual block R (nd x nlv,
// leading dimension ldr), the head coupling C (nd x nlv, ldc; may be null),
// the shifted diagonal D, the Ritz values theta and the residual norms.
// Returns a device pointer to the nd x m.ldo FP32 block, or nullptr when the
// diagonal correction is to be used.
const float *mlp_apply(Ctx &c, MlpPrecond &m, const double *D, const double *theta, const double *R, int ldr,
                       const double *C, int ldc, int nd0, const std::vector<double> &rnorm);
// Outcome of the previous iteration.  The first MLP_WARMUP iterations use the
// diagonal correction and record its residual reduction factor; a learned
// step is kept only while it reduces the residual norm (max over roots) at
// least as fast as the last diagonal step did, else the diagonal correction
// is used for the rest of the solve.
// Returns true when the last learned step is rejected by this observation.
bool mlp_observe(MlpPrecond &m, double cnx, int *used, int *rejected);
void mlp_end(Ctx &c, MlpPrecond &m);
// training-data hook, called once per iteration with the current basis
void mlp_collect_step(Ctx &c, MlpCollect &col, const double *D, const double *theta, const double *R, int ldr,
                      const double *C, int ldc, const double *B, int ldb, int nd0, int nlv,
                      const std::vector<double> &rnorm);

}  // namespace pci
