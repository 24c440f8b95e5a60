// Block Davidson for the lowest Nlv roots, entirely on the device.
//
// Reference: Diag4 (src/pconf.F90:1993-2264) with Init4, FormB0, Ortn, Mxmpy,
// FormP, Dvdsn, FormB/FormBskip (src/davidson.f90).  The iteration keeps the
// reference's structure: a 2*Nlv subspace [B | C] rebuilt every iteration,
// the Dvdsn correction (Z1 projector on rows <= Nd0, 1/(theta - D_i) on the
// other rows), the P-matrix cleanup of converged levels, the admixture stop
// and the rotation B <- [B C] P(:, 1:Nlv).
//
// Differences, all of which leave the mathematics unchanged:
//  * Ortn's one-vector-at-a-time classical Gram-Schmidt is replaced by block
//    orthonormalisation (two passes of Cholesky QR, and two passes of block
//    Gram-Schmidt of C against B).  Gram-Schmidt and Cholesky QR produce the
//    same Q (the QR factor with positive R diagonal) up to rounding.
//  * The Nlv per-column SpMVs of Mxmpy become one block SpMM.
//  * DSYEV on P is a one-workgroup parallel Jacobi in FP64; DSYEVD on Z1 is
//    rocSOLVER's dsyevd.
//
// Fast mode (PCI_DVD_FP32): H is applied in FP32 (FP32 values, FP32 vectors,
// FP32 accumulation) while the residual is above the FP32 floor, and H*B is
// carried through the rotation ([HB HC] P) instead of being recomputed, so an
// iteration costs one SpMM.  When the FP32 residual reaches the switch
// threshold, H*B is recomputed in FP64 (the refinement step) and the
// iteration continues with FP64 SpMMs; convergence is only ever declared on
// an FP64 residual.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <string>
#include <thread>

#include "internal.h"
#include "mlp.h"

namespace pci {

int eigh32_reduce(Ctx &c, rocblas_handle h, int n, float *A, float *W, Eigh32Work &w);  // eigh.hip
void eigh32_finish(Ctx &c, rocblas_handle h, int n, float *A, float *W, Eigh32Work &w);

constexpr int DVD_MAXLV = 24;            // largest Nlv handled
constexpr int DVD_MAXK = 2 * DVD_MAXLV;  // subspace width

// ------------------------------------------------------------------ kernels
// V[:, 0:nlv] = Z1[:, 0:nlv] zero-padded; other columns cleared.
// FormB0: B(:, j) = Z1(:, sel[j]) on the first nd0 rows (sel null: column j)
__global__ void k_init_B(int nd, int nd0, int nlv, int ldv, const double *__restrict__ Z, const int *__restrict__ sel,
                         double *__restrict__ V) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nd * ldv) return;
    const int r = (int)(i / ldv), j = (int)(i % ldv);
    V[i] = (j < nlv && r < nd0) ? Z[(size_t)(sel ? sel[j] : j) * nd0 + r] : 0.0;
}

// FormB0 with |Kl4| = 2: B(:, j) = the caller's start vector j (x0 vector-major)
__global__ void k_set_B(int nd, int nlv, int ldv, const double *__restrict__ x0, double *__restrict__ V) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nd * nlv) return;
    const int j = (int)(i / nd), r = (int)(i % nd);
    V[(size_t)r * ldv + j] = x0[i];
}

// J of columns j0 .. j0 + gridDim.x - 1 of Z1 (nd0 x nd0, column-major) from
// J^2 restricted to the first nd0 rows: J_av(B1, Nd0) of FormB0
// (davidson.f90:104-110; formj2.f90:355-387).  One block per column, fixed
// reduction order.
__global__ __launch_bounds__(256) void k_jav_head(int nd0, const int64_t *__restrict__ rp, const int *__restrict__ col,
                                                  const double *__restrict__ val, const double *__restrict__ Z, int j0,
                                                  double *__restrict__ out) {
    __shared__ double red[256];
    const double *z = Z + (size_t)(j0 + blockIdx.x) * nd0;
    double acc = 0.0;
    for (int n = threadIdx.x; n < nd0; n += 256)
        for (int64_t e = rp[n]; e < rp[n + 1]; e++) {
            const int k = col[e];
            if (k < nd0) {
                const double r = val[e] * z[k] * z[n];
                acc += (k != n) ? r + r : r;
            }
        }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = 0.5 * (sqrt(1.0 + red[0]) - 1.0);
}

// AvgDiag (davidson.f90:371-434): D averaged over each configuration
__global__ __launch_bounds__(256) void k_avg_diag(const int *__restrict__ conf_start, double *__restrict__ D) {
    __shared__ double red[256];
    const int a = conf_start[blockIdx.x], b = conf_start[blockIdx.x + 1];
    double acc = 0.0;
    for (int n = a + (int)threadIdx.x; n < b; n += 256) acc += D[n];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    const double avg = b > a ? red[0] / (b - a) : 0.0;
    __syncthreads();
    for (int n = a + (int)threadIdx.x; n < b; n += 256) D[n] = avg;
}

// Prj_J filter step on the unconverged columns: x <- x - f J^2 x
// (jx > 0), or the limit x <- -J^2 x of the normalised filter for jx = 0
__global__ void k_jfilter(int nd, int nlv, double *__restrict__ X, int ldx, const double *__restrict__ Y, int ldy,
                          double f, int zero_term, const int *__restrict__ conv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nd * nlv) return;
    const int64_t r = i / nlv;
    const int j = (int)(i % nlv);
    if (conv[j]) return;
    const double y = Y[r * ldy + j];
    X[r * ldx + j] = zero_term ? -y : X[r * ldx + j] - f * y;
}

__global__ __launch_bounds__(256) void k_reduce_partials(int nb, int mn, const double *__restrict__ part,
                                                        double *__restrict__ out) {
    // one block per output; fixed assignment and tree, so the sum is
    // deterministic
    __shared__ double red[256];
    const int o = blockIdx.x, tid = threadIdx.x;
    double s = 0.0;
    for (int b = tid; b < nb; b += 256) s += part[(size_t)b * mn + o];
    red[tid] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) red[tid] += red[tid + w];
        __syncthreads();
    }
    if (tid == 0) out[o] = red[0];
}

// Tall-skinny products of the Davidson basis on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64:
// D(16x16) += A(16x4) B(4x16); operand lane l holds A[l % 16][l / 16] and
// B[l / 16][l % 16], the result D[l / 16 + 4 v][l % 16] in element v).
typedef double v4d __attribute__((ext_vector_type(4)));

// C = A^T B (m x n, m <= 16 TI, n <= 16 TJ): a wavefront takes 4 rows per
// instruction (row k = l / 16 of the 4, column l % 16 of a 16-wide tile),
// so each load of a wavefront reads 4 row segments of 16 values; the 4
// wavefronts of a block interleave by 4 rows, U instructions per tile in
// flight.  Per-block partials in part (block-major m x n), summed in
// block order by k_reduce_partials (deterministic).
template <int TI, int TJ, int U>
__global__ __launch_bounds__(256) void k_tn_mfma(int nd, const double *__restrict__ A, int lda, int m,
                                                 const double *__restrict__ B, int ldb, int n, int rows_per_block,
                                                 double *__restrict__ part) {
    __shared__ double sD[4][TI * TJ * 256];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int c = lane & 15, k = lane >> 4;
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
    const int64_t r1 = min((int64_t)nd, r0 + rows_per_block);
    v4d d[TI][TJ];
#pragma unroll
    for (int ti = 0; ti < TI; ti++)
#pragma unroll
        for (int tj = 0; tj < TJ; tj++) d[ti][tj] = (v4d)(0.0);
    for (int64_t base = r0 + 4 * wv; base < r1; base += 16 * U) {  // uniform over the wavefront
        double a[U][TI], b[U][TJ];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t r = base + 16 * u + k;
            const bool ok = r < r1;
#pragma unroll
            for (int ti = 0; ti < TI; ti++) {
                const int col = 16 * ti + c;
                a[u][ti] = (ok && col < m) ? A[(size_t)r * lda + col] : 0.0;
            }
#pragma unroll
            for (int tj = 0; tj < TJ; tj++) {
                const int col = 16 * tj + c;
                b[u][tj] = (ok && col < n) ? B[(size_t)r * ldb + col] : 0.0;
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int ti = 0; ti < TI; ti++)
#pragma unroll
                for (int tj = 0; tj < TJ; tj++)
                    d[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u][ti], b[u][tj], d[ti][tj], 0, 0, 0);
    }
#pragma unroll
    for (int ti = 0; ti < TI; ti++)
#pragma unroll
        for (int tj = 0; tj < TJ; tj++)
#pragma unroll
            for (int v = 0; v < 4; v++) sD[wv][(ti * TJ + tj) * 256 + (k + 4 * v) * 16 + c] = d[ti][tj][v];
    __syncthreads();
    for (int o = threadIdx.x; o < m * n; o += 256) {
        const int i = o / n, j = o - i * n;
        const int e = ((i >> 4) * TJ + (j >> 4)) * 256 + (i & 15) * 16 + (j & 15);
        part[(size_t)blockIdx.x * m * n + o] = ((sD[0][e] + sD[1][e]) + sD[2][e]) + sD[3][e];
    }
}

// Y[:, 0:n] = beta Y + alpha X[:, 0:K] M (M: K x n row-major, n <= 16 TJ,
// K <= 48): a wavefront owns 16 rows; it loads their K columns (4 per
// instruction step: X[row0 + l % 16][k0 + l / 16]) before writing any of Y,
// so X and Y may be the same rows.
template <int TJ>
__global__ __launch_bounds__(256) void k_nn_mfma(int nd, const double *X, int ldx, int K,
                                                 const double *__restrict__ M, int n, double *Y, int ldy,
                                                 double alpha, double beta) {
    constexpr int KS = DVD_MAXK / 4;
    __shared__ double sM[DVD_MAXK * 16 * TJ];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int c = lane & 15, k = lane >> 4;
    for (int e = threadIdx.x; e < DVD_MAXK * 16 * TJ; e += 256) {
        const int kk = e / (16 * TJ), j = e - kk * (16 * TJ);
        sM[e] = (kk < K && j < n) ? M[kk * n + j] : 0.0;
    }
    __syncthreads();
    const int64_t row0 = ((int64_t)blockIdx.x * 4 + wv) * 16;
    if (row0 >= nd) return;  // uniform over the wavefront, no barrier follows
    const int64_t ra = row0 + c;
    const bool okr = ra < nd;
    double x[KS];
#pragma unroll
    for (int s = 0; s < KS; s++) {
        const int kk = 4 * s + k;
        x[s] = (okr && kk < K) ? X[(size_t)ra * ldx + kk] : 0.0;
    }
    v4d d[TJ];
#pragma unroll
    for (int tj = 0; tj < TJ; tj++) {
        d[tj] = (v4d)(0.0);
#pragma unroll
        for (int s = 0; s < KS; s++)
            if (4 * s < K)  // uniform
                d[tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[s], sM[(4 * s + k) * 16 * TJ + 16 * tj + c], d[tj],
                                                             0, 0, 0);
    }
#pragma unroll
    for (int tj = 0; tj < TJ; tj++) {
        const int j = 16 * tj + c;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int64_t r = row0 + k + 4 * v;
            if (r < nd && j < n) {
                const size_t yo = (size_t)r * ldy + j;
                Y[yo] = (beta == 0.0 ? 0.0 : beta * Y[yo]) + alpha * d[tj][v];
            }
        }
    }
}

// out[j * nd + r] = V[r, j] for the first nlv columns (vector-major output)
__global__ void k_vec_major(int nd, int nlv, const double *__restrict__ V, int ldv, double *__restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nd * nlv) return;
    const int64_t j = t / nd, r = t - j * nd;
    out[t] = V[r * ldv + j];
}

__global__ void k_copy_cols(int nd, const double *__restrict__ X, int ldx, double *__restrict__ Y, int ldy, int n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nd * n) return;
    const int64_t r = i / n;
    const int j = (int)(i % n);
    Y[r * ldy + j] = X[r * ldx + j];
}

__global__ void k_cvt_to32(int nd, const double *__restrict__ X, int ldx, float *__restrict__ Y, int ldy, int n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nd * n) return;
    const int64_t r = i / n;
    const int j = (int)(i % n);
    Y[r * ldy + j] = (float)X[r * ldx + j];
}

__global__ void k_cvt_to64(int nd, const float *__restrict__ X, int ldx, double *__restrict__ Y, int ldy, int n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nd * n) return;
    const int64_t r = i / n;
    const int j = (int)(i % n);
    Y[r * ldy + j] = (double)X[r * ldx + j];
}

// R[:, j] = HB[:, j] - theta_j B[:, j]
__global__ void k_residual(int nd, int nlv, const double *__restrict__ HB, int ldh, const double *__restrict__ B,
                           int ldb, const double *__restrict__ theta, double *__restrict__ R, int ldr) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nd * nlv) return;
    const int64_t r = i / nlv;
    const int j = (int)(i % nlv);
    R[r * ldr + j] = HB[r * ldh + j] - theta[j] * B[r * ldb + j];
}

template <typename TM>
__global__ void k_transpose(int n, const TM *__restrict__ A, TM *__restrict__ AT) {
    __shared__ TM t[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32, tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int k = ty; k < 32; k += 8)
        if (bx + tx < n && by + k < n) t[k][tx] = A[(size_t)(by + k) * n + bx + tx];
    __syncthreads();
    for (int k = ty; k < 32; k += 8)
        if (by + tx < n && bx + k < n) AT[(size_t)(bx + k) * n + by + tx] = t[tx][k];
}
template <typename TM>
void transpose_sq(int n, const TM *A, TM *AT, hipStream_t s) {
    dim3 g((unsigned)((n + 31) / 32), (unsigned)((n + 31) / 32));
    k_transpose<TM><<<g, 256, 0, s>>>(n, A, AT);
}
// Dvdsn head projector (davidson.f90:500-545), as two passes of one kernel:
//   S(j, v)  = sum_r Z1(r, n01 + j) R(r, v) / (theta_v - E1(n01 + j))
//              (0 where |theta_v - E1| < 1e-6, davidson.f90:519-531)
//   B1(r, v) = sum_j Z1(r, n01 + j) S(j, v)
// Each pass is out(c, v) = sum_{i < ni} M[i * ldm + off + c] X(i, v): one
// lane per output column c, the rows i of the sum split in slices of
// PJ_ROWS; for a slice the workgroup stages X's rows in LDS once (every lane
// reads the same X value) and streams the slice's rows of M, which are
// contiguous in c (the first pass reads Z1^T, a transposed copy made once,
// the second Z1).  M may hold FP32 eigenvectors (fast mode): the sums are
// FP64 either way.  Partial sums per slice go to part; k_proj_sum adds them
// in slice order and applies the 1/(theta - E) scaling of the first pass.
constexpr int PJ_ROWS = 128;
template <typename TM, int NVB>
__global__ __launch_bounds__(256) void k_projt(int ni, int nc, int nv, const TM *__restrict__ M, int64_t ldm,
                                               int off, const double *__restrict__ X, int ldxr, int ldxc,
                                               double *__restrict__ part) {
    __shared__ double sx[PJ_ROWS * NVB];
    const int c = blockIdx.x * 256 + threadIdx.x, sl = blockIdx.y;
    const int i0 = sl * PJ_ROWS, nr = min(PJ_ROWS, ni - i0);
    for (int e = threadIdx.x; e < nr * NVB; e += 256) {
        const int r = e / NVB, v = e - r * NVB;
        sx[e] = v < nv ? X[(size_t)(i0 + r) * ldxr + (size_t)v * ldxc] : 0.0;
    }
    __syncthreads();
    if (c >= nc) return;
    double acc[NVB];
#pragma unroll
    for (int v = 0; v < NVB; v++) acc[v] = 0.0;
    const TM *mc = M + (size_t)i0 * ldm + off + c;
    for (int r = 0; r < nr; r++) {
        const double m = (double)mc[(size_t)r * ldm];
#pragma unroll
        for (int v = 0; v < NVB; v++) acc[v] += m * sx[r * NVB + v];
    }
    double *o = part + ((size_t)sl * nc + c) * NVB;
#pragma unroll
    for (int v = 0; v < NVB; v++) o[v] = acc[v];
}

template <int NVB>
__global__ void k_proj_sum(int nsl, int nc, int nv, const double *__restrict__ part, double *__restrict__ out, int ldoc,
                           int ldov, const double *__restrict__ theta, const double *__restrict__ E, int e0) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nc * nv) return;
    const int c = (int)(t / nv), v = (int)(t - (int64_t)c * nv);
    double acc = 0.0;
    for (int sl = 0; sl < nsl; sl++) acc += part[((size_t)sl * nc + c) * NVB + v];
    if (theta) {
        const double dn = theta[v] - E[e0 + c];
        acc = fabs(dn) < 1e-6 ? 0.0 : acc / dn;
    }
    out[(size_t)c * ldoc + (size_t)v * ldov] = acc;
}

template <typename TM>
void projt_pass(int ni, int nc, int nv, const TM *M, int64_t ldm, int off, const double *X, int ldxr, int ldxc,
                double *part, double *out, int ldoc, int ldov, const double *theta, const double *E, int e0,
                hipStream_t s) {
    const int nsl = (ni + PJ_ROWS - 1) / PJ_ROWS;
    const dim3 g((unsigned)((nc + 255) / 256), (unsigned)nsl);
    const unsigned gs = (unsigned)(((int64_t)nc * nv + 255) / 256);
#define PCI_PROJT(NVB)                                                                                  \
    do {                                                                                                \
        k_projt<TM, NVB><<<g, 256, 0, s>>>(ni, nc, nv, M, ldm, off, X, ldxr, ldxc, part);               \
        k_proj_sum<NVB><<<gs, 256, 0, s>>>(nsl, nc, nv, part, out, ldoc, ldov, theta, E, e0);           \
    } while (0)
    if (nv <= 4) PCI_PROJT(4);
    else if (nv <= 8) PCI_PROJT(8);
    else if (nv <= 12) PCI_PROJT(12);
    else if (nv <= 16) PCI_PROJT(16);
    else PCI_PROJT(24);
#undef PCI_PROJT
}

// Correction: rows < nd0 from the Z1 projector (B1, column-major nd0 x nlv),
// rows >= nd0 either r_i / (theta - D_i) or the MLP output. Converged -> 0.
// R and C may be the same block (each element is read, then overwritten, by
// one thread).  Rows >= nd0: learned correction if given, else
// (r + cpl) / (theta - D) with the head-to-tail coupling cpl (may be null:
// pconf's diagonal correction r / (theta - D)).
__global__ void k_correction(int nd, int nd0, int nlv, const double *R, int ldr,
                             const double *__restrict__ B1, const double *__restrict__ D,
                             const double *__restrict__ theta, const int *__restrict__ conv,
                             const float *__restrict__ learned, int ldl, const double *__restrict__ cpl, int ldcp,
                             double *C, int ldc) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nd * nlv) return;
    const int64_t r = i / nlv;
    const int j = (int)(i % nlv);
    double c;
    if (conv[j]) c = 0.0;
    else if (r < nd0) c = B1[(size_t)j * nd0 + r];
    else if (learned) c = (double)learned[r * ldl + j];
    else c = (R[r * ldr + j] + (cpl ? cpl[r * ldcp + j] : 0.0)) / (theta[j] - D[r]);
    C[r * ldc + j] = c;
}

// T[r * ldt + j] = B1[j * nd0 + r]: the head correction (column-major) as a
// row-major block for the coupling gather
__global__ void k_head_rows(int nd0, int nlv, const double *__restrict__ B1, double *__restrict__ T, int ldt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nd0 * nlv) return;
    const int64_t r = i / nlv;
    const int j = (int)(i % nlv);
    T[r * ldt + j] = B1[(size_t)j * nd0 + r];
}

// theta_j = P11(j, j); convergence flags ||r_j|| < Crt4 from the Gram matrix
// of the residuals (the host takes the same decision from the same values)
__global__ void k_theta_conv(int nlv, const double *__restrict__ P11, const double *__restrict__ G, double crt4,
                             double *__restrict__ theta, int *__restrict__ conv) {
    const int j = threadIdx.x;
    if (j >= nlv) return;
    theta[j] = P11[j * nlv + j];
    if (G) conv[j] = (std::sqrt(std::fabs(G[j * nlv + j])) < crt4) ? 1 : 0;
}
// columns scaled to unit norm from the diagonal of their Gram matrix G
__global__ void k_scale_gram(int nd, int n, double *__restrict__ V, int ldv, const double *__restrict__ G) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nd * n) return;
    const int64_t r = i / n;
    const int j = (int)(i % n);
    const double g = G[j * n + j];
    V[r * ldv + j] *= g > 0.0 ? 1.0 / std::sqrt(g) : 0.0;
}
__global__ void k_scale_cols(int nd, int n, double *__restrict__ V, int ldv, const double *__restrict__ s) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nd * n) return;
    const int64_t r = i / n;
    const int j = (int)(i % n);
    V[r * ldv + j] *= s[j];
}

// Cholesky of the m x m Gram matrix G (row-major) and inverse of its upper
// factor: Rinv such that V * Rinv has orthonormal columns.  Columns flagged
// in skip (converged corrections, all zero) keep Rinv(j,j) = 1 and are
// decoupled from the others.  A pivot below 1e-9 (critN of Ortn,
// davidson.f90:717) is a failure.  One wavefront, m <= DVD_MAXK, in LDS:
// right-looking column steps, then one column of Rinv = (L^T)^-1 per lane.
__global__ __launch_bounds__(64) void k_chol_inv(int m, const double *__restrict__ G, const int *__restrict__ skip,
                                                 double *__restrict__ Rinv, int *__restrict__ fail) {
    __shared__ double L[DVD_MAXK][DVD_MAXK + 1];
    __shared__ int sk[DVD_MAXK];
    __shared__ int bad;
    const int tid = threadIdx.x;
    for (int e = tid; e < m * m; e += 64) L[e / m][e % m] = G[e];
    for (int i = tid; i < m; i += 64) sk[i] = skip ? skip[i] : 0;
    if (tid == 0) bad = 0;
    __syncthreads();
    for (int j = 0; j < m; j++) {
        if (tid == 0) {
            if (sk[j]) {
                L[j][j] = 1.0;
            } else {
                double d = L[j][j];
                if (!(d >= 1e-9)) { bad = 1; d = 1.0; }
                L[j][j] = sqrt(d);
            }
        }
        __syncthreads();
        for (int i = j + 1 + tid; i < m; i += 64) L[i][j] = (sk[j] || sk[i]) ? 0.0 : L[i][j] / L[j][j];
        __syncthreads();
        // trailing update of the lower triangle: L[i][k] -= L[i][j] L[k][j], j < k <= i
        const int t = m - j - 1;
        for (int e = tid; e < t * t; e += 64) {
            const int i = j + 1 + e / t, k = j + 1 + e % t;
            if (k <= i) L[i][k] -= L[i][j] * L[k][j];
        }
        __syncthreads();
    }
    // Rinv(:, c): solve L^T x = e_c (upper triangular), rows c down to 0
    for (int c = tid; c < m; c += 64) {
        for (int i = m - 1; i > c; i--) Rinv[i * m + c] = 0.0;
        double x[DVD_MAXK];
#pragma unroll
        for (int i = 0; i < DVD_MAXK; i++) x[i] = 0.0;
        for (int i = c; i >= 0; i--) {
            double s = (i == c) ? 1.0 : 0.0;
#pragma unroll
            for (int k = 0; k < DVD_MAXK; k++)
                if (k > i && k <= c) s -= L[k][i] * x[k];
            const double v = s / L[i][i];
#pragma unroll
            for (int k = 0; k < DVD_MAXK; k++)
                if (k == i) x[k] = v;
            Rinv[i * m + c] = v;
        }
    }
    if (tid == 0 && bad) *fail = 1;
}

// P (2nlv x 2nlv, row-major) from P11 = B^T HB (nlv x nlv) and
// P2 = [B C]^T HC (2nlv x nlv), lower parts mirrored as FormP does
// (davidson.f90:326-366), converged levels cleaned.
__global__ __launch_bounds__(256) void k_assemble_P(int nlv, const double *__restrict__ P11,
                                                    const double *__restrict__ P2, const int *__restrict__ conv,
                                                    double vmax, double *__restrict__ P) {
    // one thread per element (r, c): the blocks of P are the lower triangle
    // of P11, the rows of P2 and their mirror images
    const int n2 = 2 * nlv;
    for (int e = threadIdx.x; e < n2 * n2; e += blockDim.x) {
        const int r = e / n2, c = e - r * n2;
        const int hi = r > c ? r : c, lo = r > c ? c : r;
        double v;
        if (hi < nlv) v = P11[hi * nlv + lo];
        else if (lo < nlv) v = P2[lo * nlv + (hi - nlv)];
        else v = P2[hi * nlv + (lo - nlv)];
        // converged level i: row and column Nlv + i cleared, diagonal moved above the spectrum
        if (r >= nlv && conv[r - nlv]) v = (r == c) ? 2.0 * vmax - P11[(r - nlv) * nlv + (r - nlv)] + 1.0 : 0.0;
        else if (c >= nlv && conv[c - nlv]) v = 0.0;
        P[e] = v;
    }
}

// Symmetric eigen-decomposition of the small n x n projected matrix P
// (n = 2 Nlv <= 48, row-major) on the host by cyclic Jacobi rotations, in
// place of DSYEV (pconf.F90:2130-2160).  Out: E ascending, Q row-major with
// column j the eigenvector of E[j].  A 48 x 48 solve is a few tens of
// microseconds on one core, less than one round trip through a kernel.
static void jacobi_host(int n, std::vector<double> A, std::vector<double> &E, std::vector<double> &Q) {
    std::vector<double> V((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) V[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0.0, dia = 0.0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) (i == j ? dia : off) += A[(size_t)i * n + j] * A[(size_t)i * n + j];
        if (off == 0.0 || off <= 1e-32 * dia) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                const double apq = A[(size_t)p * n + q];
                if (apq == 0.0) continue;
                const double tau = (A[(size_t)q * n + q] - A[(size_t)p * n + p]) / (2.0 * apq);
                const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                const double c = 1.0 / std::sqrt(1.0 + t * t), sn = t * c;
                for (int k = 0; k < n; k++) {  // A <- A J
                    const double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
                    A[(size_t)k * n + p] = c * akp - sn * akq;
                    A[(size_t)k * n + q] = sn * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {  // A <- J^T A
                    const double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
                    A[(size_t)p * n + k] = c * apk - sn * aqk;
                    A[(size_t)q * n + k] = sn * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {  // V <- V J
                    const double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
                    V[(size_t)k * n + p] = c * vkp - sn * vkq;
                    V[(size_t)k * n + q] = sn * vkp + c * vkq;
                }
            }
    }
    std::vector<int> idx(n);
    for (int i = 0; i < n; i++) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int x, int y) { return A[(size_t)x * n + x] < A[(size_t)y * n + y]; });
    E.resize(n);
    Q.resize((size_t)n * n);
    for (int j = 0; j < n; j++) {
        E[j] = A[(size_t)idx[j] * n + idx[j]];
        for (int i = 0; i < n; i++) Q[(size_t)i * n + j] = V[(size_t)i * n + idx[j]];
    }
}

// The same P as k_assemble_P with vmax taken on the device: max of the
// stored vmax and the current Ritz values (the host form's
// vmax = max(vmax, max theta) before FormP), so the iteration needs no host
// value.
__global__ __launch_bounds__(256) void k_assemble_P_dev(int nlv, const double *__restrict__ P11,
                                                        const double *__restrict__ P2, const int *__restrict__ conv,
                                                        const double *__restrict__ theta, const double *__restrict__ vm,
                                                        double *__restrict__ P) {
    double vmax = vm[0];
    for (int j = 0; j < nlv; j++) vmax = theta[j] > vmax ? theta[j] : vmax;
    const int n2 = 2 * nlv;
    for (int e = threadIdx.x; e < n2 * n2; e += blockDim.x) {
        const int r = e / n2, c = e - r * n2;
        const int hi = r > c ? r : c, lo = r > c ? c : r;
        double v;
        if (hi < nlv) v = P11[hi * nlv + lo];
        else if (lo < nlv) v = P2[lo * nlv + (hi - nlv)];
        else v = P2[hi * nlv + (lo - nlv)];
        if (r >= nlv && conv[r - nlv]) v = (r == c) ? 2.0 * vmax - P11[(r - nlv) * nlv + (r - nlv)] + 1.0 : 0.0;
        else if (c >= nlv && conv[c - nlv]) v = 0.0;
        P[e] = v;
    }
}

// Round r (0 .. m-2) of the round-robin pairing of m (even) indices, pair k
// (0 .. m/2-1): index m-1 meets r, the others meet (r + k, r - k) mod (m-1).
// Over the m-1 rounds every pair of indices meets exactly once, and within a
// round the pairs are disjoint.
__host__ __device__ inline void rr_pair(int m, int r, int k, int &p, int &q) {
    if (k == 0) {
        p = r;
        q = m - 1;
    } else {
        p = (r + k) % (m - 1);
        q = (r - k + (m - 1)) % (m - 1);
    }
}

// Eigen-decomposition of P on the device by Jacobi rotations in parallel
// order: per round the n/2 disjoint rotations of the round-robin pairing
// (rr_pair; an odd n gets a dummy index that never rotates) are computed from
// the current matrix and applied together (columns of A and V, then rows of
// A), by a 256-thread workgroup with P in LDS; a sweep is the m - 1 rounds,
// so every off-diagonal pair is rotated once per sweep as in jacobi_host's
// cyclic order.  Same rotation formula and stopping rule as jacobi_host.
// Out: E (n, ascending), Q (n x nlv row-major: the first nlv eigenvectors,
// the FormB rotation), vm[0] = max of the first nlv eigenvalues.
__global__ __launch_bounds__(256) void k_jacobi_dev(int n, int nlv, const double *__restrict__ P,
                                                    double *__restrict__ Q, double *__restrict__ E,
                                                    double *__restrict__ vm) {
    __shared__ double A[DVD_MAXK * DVD_MAXK], V[DVD_MAXK * DVD_MAXK];
    __shared__ double cs[DVD_MAXK / 2][2];
    __shared__ int pq[DVD_MAXK / 2][2];
    __shared__ double red[2][4];
    __shared__ int idx[DVD_MAXK];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int m = n + (n & 1), half = m / 2;
    for (int e = t; e < n * n; e += 256) {
        A[e] = P[e];
        V[e] = (e / n == e % n) ? 1.0 : 0.0;
    }
    __syncthreads();
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0.0, dia = 0.0;
        for (int e = t; e < n * n; e += 256) {
            const double a = A[e];
            if (e / n == e % n) dia += a * a;
            else off += a * a;
        }
        for (int o = 32; o > 0; o >>= 1) {
            off += __shfl_xor(off, o);
            dia += __shfl_xor(dia, o);
        }
        if (lane == 0) {
            red[0][wv] = off;
            red[1][wv] = dia;
        }
        __syncthreads();
        off = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        dia = red[1][0] + red[1][1] + red[1][2] + red[1][3];
        if (off == 0.0 || off <= 1e-32 * dia) break;  // uniform
        for (int r = 0; r < m - 1; r++) {
            if (t < half) {  // the round's rotations, from the current A
                int p, q;
                rr_pair(m, r, t, p, q);
                double c = 1.0, sn = 0.0;
                if (p < n && q < n) {
                    if (p > q) {
                        const int x = p;
                        p = q;
                        q = x;
                    }
                    const double apq = A[p * n + q];
                    if (apq != 0.0) {
                        const double tau = (A[q * n + q] - A[p * n + p]) / (2.0 * apq);
                        const double tt = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                        c = 1.0 / sqrt(1.0 + tt * tt);
                        sn = tt * c;
                    }
                } else {
                    p = q = 0;  // the dummy index's pair: identity
                }
                pq[t][0] = p;
                pq[t][1] = q;
                cs[t][0] = c;
                cs[t][1] = sn;
            }
            __syncthreads();
            for (int e = t; e < 2 * half * n; e += 256) {  // A <- A J, V <- V J
                const int w = e / (half * n), f = e - w * half * n;
                const int k = f / n, i = f - k * n;
                const double c = cs[k][0], sn = cs[k][1];
                if (sn == 0.0 && c == 1.0) continue;
                const int p = pq[k][0], q = pq[k][1];
                double *M = w ? V : A;
                const double mp = M[i * n + p], mq = M[i * n + q];
                M[i * n + p] = c * mp - sn * mq;
                M[i * n + q] = sn * mp + c * mq;
            }
            __syncthreads();
            for (int f = t; f < half * n; f += 256) {  // A <- J^T A
                const int k = f / n, i = f - k * n;
                const double c = cs[k][0], sn = cs[k][1];
                if (sn == 0.0 && c == 1.0) continue;
                const int p = pq[k][0], q = pq[k][1];
                const double ap = A[p * n + i], aq = A[q * n + i];
                A[p * n + i] = c * ap - sn * aq;
                A[q * n + i] = sn * ap + c * aq;
            }
            __syncthreads();
        }
    }
    if (t == 0) {  // stable ascending order of the diagonal
        for (int i = 0; i < n; i++) idx[i] = i;
        for (int i = 1; i < n; i++) {
            const int x = idx[i];
            int j = i - 1;
            while (j >= 0 && A[idx[j] * n + idx[j]] > A[x * n + x]) {
                idx[j + 1] = idx[j];
                j--;
            }
            idx[j + 1] = x;
        }
    }
    __syncthreads();
    for (int j = t; j < n; j += 256) E[j] = A[idx[j] * n + idx[j]];
    for (int e = t; e < n * nlv; e += 256) {
        const int i = e / nlv, j = e - i * nlv;
        Q[e] = V[i * n + idx[j]];
    }
    if (t == 0) {
        double mx = -1e10;
        for (int j = 0; j < nlv; j++) {
            const double ej = A[idx[j] * n + idx[j]];
            mx = ej > mx ? ej : mx;
        }
        vm[0] = mx;
    }
}

// <x|J^2|x> over the stored lower triangle (J_av, formj2.f90:355-387)
__global__ __launch_bounds__(256) void k_jav(int nd, const int64_t *__restrict__ rp, const int *__restrict__ col,
                                             const double *__restrict__ val, const double *__restrict__ V, int ldv,
                                             int nlv, double *__restrict__ part) {
    // one thread per row (J^2 rows are short: one configuration block); the
    // block's sums go to part[blockIdx.x * nlv + v], reduced in block order
    __shared__ double red[256];
    const int row = blockIdx.x * 256 + threadIdx.x;
    double acc[DVD_MAXLV];
#pragma unroll
    for (int v = 0; v < DVD_MAXLV; v++) acc[v] = 0.0;
    if (row < nd) {
        for (int64_t j = rp[row]; j < rp[row + 1]; j++) {
            const int k = col[j];
            const double t = val[j];
#pragma unroll
            for (int v = 0; v < DVD_MAXLV; v++)
                if (v < nlv) {
                    double r = t * V[(size_t)k * ldv + v] * V[(size_t)row * ldv + v];
                    if (k != row) r = r + r;
                    acc[v] += r;
                }
        }
    }
#pragma unroll
    for (int v = 0; v < DVD_MAXLV; v++) {
        if (v >= nlv) break;
        red[threadIdx.x] = acc[v];
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(size_t)blockIdx.x * nlv + v] = red[0];
        __syncthreads();
    }
}

// ------------------------------------------------------------------ driver
namespace {

inline unsigned blocks_for(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

struct Dvd {
    Ctx &c;
    hipStream_t s;
    int nd, nlv, nd0, ldv;
    DBuf<double> V, HV, D, Z1, E1, S, B1, part, small, P, norms, theta;
    DBuf<double> eigs, vm;  // device eigen-decomposition of P: eigenvalues, vmax (current | snapshot)
    DBuf<double> projpart;  // head projector partial sums (k_projt)
    DBuf<int> conv, fail;
    DBuf<float> X32, Y32;
    int ld32 = 0;
    DBuf<double> Z1T;       // Z1 transposed (plain mode head projector)
    int eigh_path = -1;     // fast mode: 1 cooperative tridiagonalisation, 0 ssyevd
    DBuf<float> Z32, Z32T;  // fast mode: FP32 head eigenvectors and their transpose
    DBuf<double> X64;   // compact copy of the SpMM input block (rows 16-byte aligned)
    int ld64 = 0;
    DBuf<double> PJx, PJy;  // J-projection work blocks (nd x ld64)
    DBuf<double> SnapB;     // basis before a learned step (nd x nlv)
    DBuf<double> SnapHB;    // fast mode: H*B at that point (nd x nlv)
    // head-to-tail coupling of the correction (PCI_DVD_COUPLED)
    DBuf<double> Tb, Cp;
    DBuf<float> Tb32, Cp32;
    int ldt = 0, ldt32 = 0;
    std::vector<double> h_norm, h_theta, h_E;
    std::vector<int> h_conv;
    rocblas_handle blas = nullptr;
    std::vector<hipEvent_t> evs;  // start/stop pairs of every SpMM, read after the solve
    int nev = 0;
    std::vector<char> ev32;  // per event pair: an FP32 SpMM
    double t_spmm = 0.0, t_spmm32 = 0.0, bytes = 0.0;
    int n32 = 0, n64 = 0;
    int max_parts = 0;

    explicit Dvd(Ctx &cc) : c(cc), s(cc.stream) {
        BlockPool *bp = &cc.pool;  // the solve's buffers come from the context's pool
        for (DBuf<double> *b : {&V, &HV, &D, &Z1, &E1, &S, &B1, &part, &small, &P, &norms, &theta, &eigs, &vm, &projpart, &Z1T,
                                &X64, &PJx, &PJy, &SnapB, &SnapHB, &Tb, &Cp})
            b->pool = bp;
        for (DBuf<int> *b : {&conv, &fail}) b->pool = bp;
        for (DBuf<float> *b : {&X32, &Y32, &Z32, &Z32T, &Tb32, &Cp32}) b->pool = bp;
    }
    ~Dvd() {
        for (hipEvent_t e : evs) (void)hipEventDestroy(e);
    }

    // C(m x n) = A^T B over all rows, result row-major in out (device); on
    // the FP64 matrix cores (k_tn_mfma), m <= 48, n <= 32
    void tn(const double *A, int lda, int m, const double *B, int ldb, int n, double *out) {
        PCI_REQUIRE(m >= 1 && n >= 1 && m <= 48 && n <= 32, PCI_E_UNSUPPORTED, "block too wide");
        constexpr int U = 4;
        // about 1024 blocks (the partials buffer), whole steps of 16 U rows each
        int nb = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (nd + 16 * U - 1) / (16 * U)));
        int rows = (int)((nd + nb - 1) / nb);
        rows = (rows + 16 * U - 1) / (16 * U) * (16 * U);
        nb = (int)((nd + rows - 1) / rows);
        const int ti = (m + 15) / 16, tj = (n + 15) / 16;
#define PCI_TN(I, J) \
        if (ti == I && tj == J) k_tn_mfma<I, J, U><<<nb, 256, 0, s>>>(nd, A, lda, m, B, ldb, n, rows, part.p);
        PCI_TN(1, 1) PCI_TN(1, 2) PCI_TN(2, 1) PCI_TN(2, 2) PCI_TN(3, 1) PCI_TN(3, 2)
#undef PCI_TN
        k_reduce_partials<<<m * n, 256, 0, s>>>(nb, m * n, part.p, out);
    }
    // Y = alpha X[:, 0:K] M + beta Y (k_nn_mfma)
    void nn(const double *X, int ldx, int K, const double *M, int n, double *Y, int ldy, double alpha, double beta) {
        PCI_REQUIRE(K <= DVD_MAXK && n <= DVD_MAXLV, PCI_E_UNSUPPORTED, "block too wide");
        const unsigned g = (unsigned)((nd + 63) / 64);
        if (n <= 16) k_nn_mfma<1><<<g, 256, 0, s>>>(nd, X, ldx, K, M, n, Y, ldy, alpha, beta);
        else k_nn_mfma<2><<<g, 256, 0, s>>>(nd, X, ldx, K, M, n, Y, ldy, alpha, beta);
    }
    hipEvent_t next_event() {
        if (nev == (int)evs.size()) {
            hipEvent_t e;
            PCI_HIP(hipEventCreate(&e));
            evs.push_back(e);
        }
        return evs[nev++];
    }
    // sum of the SpMM times (waits for the stream); t_spmm32 the FP32 part
    void collect_spmm_time() {
        PCI_HIP(hipStreamSynchronize(s));
        t_spmm = t_spmm32 = 0.0;
        for (int i = 0; i + 1 < nev; i += 2) {
            float ms = 0.f;
            PCI_HIP(hipEventElapsedTime(&ms, evs[i], evs[i + 1]));
            t_spmm += ms;
            if (ev32[i / 2]) t_spmm32 += ms;
        }
    }
    // H X for nlv columns; the events bracket the SpMM launches alone (not
    // the precision conversions around them)
    void spmm(bool fp32, const double *X, int ldx, double *Y, int ldy) {
        if (fp32) {
            k_cvt_to32<<<blocks_for((int64_t)nd * nlv), 256, 0, s>>>(nd, X, ldx, X32.p, ld32, nlv);
            PCI_HIP(hipEventRecord(next_event(), s));
            spmm_f32(c, X32.p, ld32, Y32.p, ld32, nlv, s, true);
            PCI_HIP(hipEventRecord(next_event(), s));
            k_cvt_to64<<<blocks_for((int64_t)nd * nlv), 256, 0, s>>>(nd, Y32.p, ld32, Y, ldy, nlv);
            n32++;
        } else {
            // gather from a compact nd x ld64 copy: a third of the footprint
            // of the 3 Nlv-wide basis array, so more of it stays in L2
            k_copy_cols<<<blocks_for((int64_t)nd * nlv), 256, 0, s>>>(nd, X, ldx, X64.p, ld64, nlv);
            PCI_HIP(hipEventRecord(next_event(), s));
            spmm_f64(c, X64.p, ld64, Y, ldy, nlv, s, true);
            PCI_HIP(hipEventRecord(next_event(), s));
            n64++;
        }
        ev32.push_back(fp32 ? 1 : 0);
        bytes += spmm_bytes(c, nlv, fp32 ? 32 : 64);
    }
    // orthonormalise columns [c0, c0+n) of V (two Cholesky-QR passes);
    // also applies the same transform to columns [h0, h0+n) of HV if h0 >= 0.
    // passes: 2 for fresh corrections (Cholesky-QR2); 1 for the basis B,
    // which the rotation [B C] Q keeps orthonormal to rounding
    void cholqr(int c0, int n, const int *skip_dev, int h0, int passes = 2) {
        double *G = small.p, *Rinv = small.p + 64 * 64;
        for (int pass = 0; pass < passes; pass++) {
            tn(V.p + c0, ldv, n, V.p + c0, ldv, n, G);
            k_chol_inv<<<1, 64, 0, s>>>(n, G, skip_dev, Rinv, fail.p);
            nn(V.p + c0, ldv, n, Rinv, n, V.p + c0, ldv, 1.0, 0.0);
            if (h0 >= 0) nn(HV.p + h0, 2 * nlv, n, Rinv, n, HV.p + h0, 2 * nlv, 1.0, 0.0);
        }
    }
};

}  // namespace

void BatchGate::arrive(Ctx &c, bool wait) {
    std::unique_lock<std::mutex> lk(m);
    if (!c.gate_arrived) {
        c.gate_arrived = true;
        arrived++;
        cv.notify_all();
    }
    if (wait) cv.wait(lk, [&] { return arrived >= n; });
}

void BatchGate::finish(Ctx &c, bool wait) {
    std::unique_lock<std::mutex> lk(m);
    if (!c.gate_finished) {
        c.gate_finished = true;
        finished++;
        cv.notify_all();
    }
    if (wait) cv.wait(lk, [&] { return finished >= n; });
}

// Exclusive use of the GPU for the head eigensolve of a batched solve (no-op
// outside a batch).  The destructor releases the GPU on an error path.
struct GateScope {
    Ctx &c;
    bool held = false;
    GateScope(Ctx &c_, hipStream_t s) : c(c_) {
        if (!c.gate) return;
        PCI_HIP(hipStreamSynchronize(s));
        c.gate->arrive(c, true);
        c.gate->excl.lock();
        held = true;
    }
    void leave(hipStream_t s) {
        if (!held) return;
        const hipError_t e = hipStreamSynchronize(s);
        c.gate->excl.unlock();
        held = false;
        c.gate->finish(c, true);
        PCI_HIP(e);
    }
    ~GateScope() {
        if (!held) return;
        c.gate->excl.unlock();
        c.gate->finish(c, false);
    }
};

// pconf's restart and checkpoint around Diag4 (pci_gpu_davidson_ex)
struct DvdRestart {
    const double *tk0 = nullptr, *x0 = nullptr;  // start energies and vectors (nlv of them)
    int every = 0;                               // checkpoint period in iterations (0: none)
    pci_dvd_checkpoint_fn cb = nullptr;
    void *user = nullptr;
};

static void davidson(Ctx &c, const pci_dvd_params &prm, double *tk, double *tj, double *vecs, pci_dvd_stats *st,
                     MlpCollect *col = nullptr, const DvdRestart *rs = nullptr) {
    Matrix &H = c.H;
    PCI_REQUIRE(H.valid, PCI_E_ARG, "build or load H before pci_gpu_davidson");
    const int nd = H.nd, nlv = prm.nlv;
    PCI_REQUIRE(nlv >= 1 && nlv <= DVD_MAXLV && nlv <= nd, PCI_E_ARG, "Nlv out of range (1..24, <= Nd)");
    int nd0 = prm.nd0;
    PCI_REQUIRE(nd0 >= nlv && nd0 <= nd, PCI_E_ARG, "Nd0 must satisfy Nlv <= Nd0 <= Nd");
    PCI_REQUIRE(nd0 <= 20000, PCI_E_UNSUPPORTED, "Nd0 above 20000");
    PCI_REQUIRE((prm.flags & ~(PCI_DVD_FP32 | PCI_DVD_LEARNED | PCI_DVD_TRACE | PCI_DVD_COUPLED | PCI_DVD_JSEL |
                               PCI_DVD_NOADMIX)) == 0,
                PCI_E_ARG, "unknown Davidson flag");
    // A head block that is the whole matrix (Nd0 = Nd, as in small CSF
    // spaces) is solved exactly by Init4, and pconf's Davidson stops at its
    // first residual: the solve then runs in FP64 with the plain correction
    // (an FP32 head would leave FP32-accurate start vectors whose corrections
    // lie in the space already spanned).
    const bool whole_head = (nd0 == nd);
    const bool fast = (prm.flags & PCI_DVD_FP32) != 0 && !whole_head;
    const bool learned = (prm.flags & PCI_DVD_LEARNED) != 0 && !whole_head;
    const bool trace = (prm.flags & PCI_DVD_TRACE) != 0;
    const bool coupled = (prm.flags & PCI_DVD_COUPLED) != 0 && !whole_head;
    const bool jsel = (prm.flags & PCI_DVD_JSEL) != 0;
    const bool admix_test = (prm.flags & PCI_DVD_NOADMIX) == 0;
    PCI_REQUIRE(!learned || (c.have_space && c.nd == H.nd && !H.csf), PCI_E_ARG,
                "the learned correction needs H in the determinant basis");
    if (jsel) {
        PCI_REQUIRE(c.J.valid && !c.jt.empty(), PCI_E_ARG,
                    "J selection needs J^2 (pci_gpu_build_j) and the J terms (pci_gpu_set_jterms)");
        PCI_REQUIRE(c.have_space && c.nd == H.nd && !H.csf && c.J.nd == H.nd, PCI_E_ARG,
                    "J selection needs H in the determinant basis");
        ensure_jfull(c);
    }
    const double crt4 = prm.crt4;
    auto wall0 = std::chrono::steady_clock::now();

    Dvd d(c);
    hipStream_t s = c.stream;
    d.nd = nd; d.nlv = nlv; d.nd0 = nd0; d.ldv = 3 * nlv;
    if (!c.blas) {
        rocblas_handle h = nullptr;
        PCI_REQUIRE(rocblas_create_handle(&h) == rocblas_status_success, PCI_E_HIP, "rocblas_create_handle failed");
        c.blas = h;
    }
    d.blas = (rocblas_handle)c.blas;  // owned by the context
    rocblas_set_stream(d.blas, s);
    d.V.alloc((size_t)nd * d.ldv);
    // pinned host staging: P11 | G | P | Q block | fail flag | eigenvalues of P
    const size_t pin_need = 4 * 64 * 64 + 8 + 64;
    if (c.pin_n < pin_need) {
        if (c.pin) PCI_HIP(hipHostFree(c.pin));
        c.pin = nullptr;
        c.pin_n = 0;
        PCI_HIP(hipHostMalloc((void **)&c.pin, pin_need * sizeof(double), hipHostMallocDefault));
        c.pin_n = pin_need;
    }
    double *pin_P11 = c.pin, *pin_G = c.pin + 64 * 64, *pin_P = c.pin + 2 * 64 * 64, *pin_Q = c.pin + 3 * 64 * 64;
    int *pin_fail = (int *)(c.pin + 4 * 64 * 64);
    double *pin_E = c.pin + 4 * 64 * 64 + 8;
    d.ld64 = (nlv + 1) / 2 * 2;
    d.X64.alloc((size_t)nd * d.ld64);
    PCI_HIP(hipMemsetAsync(d.X64.p, 0, (size_t)nd * d.ld64 * sizeof(double), s));
    if (jsel) {
        d.PJx.alloc((size_t)nd * d.ld64);
        d.PJy.alloc((size_t)nd * d.ld64);
        PCI_HIP(hipMemsetAsync(d.PJx.p, 0, (size_t)nd * d.ld64 * sizeof(double), s));
    }
    if (coupled && nd0 < nd) {
        head_coupling_setup(c, nd0, s);
        d.ldt = (nlv + 1) / 2 * 2;
        d.Tb.alloc((size_t)nd0 * d.ldt);
        d.Cp.alloc((size_t)nd * nlv);
        PCI_HIP(hipMemsetAsync(d.Tb.p, 0, (size_t)nd0 * d.ldt * sizeof(double), s));
        PCI_HIP(hipMemsetAsync(d.Cp.p, 0, (size_t)nd * nlv * sizeof(double), s));
        if (fast) {
            d.ldt32 = (nlv + 3) / 4 * 4;
            d.Tb32.alloc((size_t)nd0 * d.ldt32);
            d.Cp32.alloc((size_t)nd * nlv);
            PCI_HIP(hipMemsetAsync(d.Tb32.p, 0, (size_t)nd0 * d.ldt32 * sizeof(float), s));
            PCI_HIP(hipMemsetAsync(d.Cp32.p, 0, (size_t)nd * nlv * sizeof(float), s));
        }
    }
    d.D.alloc(nd);
    d.Z1.alloc((size_t)nd0 * nd0);
    d.E1.alloc(nd0);
    d.S.alloc((size_t)nd0 * nlv);
    d.B1.alloc((size_t)nd0 * nlv);
    d.part.alloc(1024 * 2048);
    d.small.alloc(4 * 64 * 64);
    d.P.alloc(64 * 64);
    d.norms.alloc(64 * 64);
    d.theta.alloc(64);
    d.eigs.alloc(64);
    d.vm.alloc(4);
    d.conv.alloc(64);
    d.fail.alloc(4);
    if (fast) {
        d.HV.alloc((size_t)nd * 2 * nlv);
        ensure_fp32(c);
        d.ld32 = (nlv + 3) / 4 * 4;
        d.X32.alloc((size_t)nd * d.ld32);
        d.Y32.alloc((size_t)nd * d.ld32);
        PCI_HIP(hipMemsetAsync(d.X32.p, 0, (size_t)nd * d.ld32 * sizeof(float), s));
    }
    MlpPrecond mlp;
    bool snap_valid = false;
    // learned steps are provisional until LRN_COMMIT of them in a row have
    // each reduced the residual as fast as the diagonal step did; a rejected
    // step returns the solve to the state before the first of them
    constexpr int LRN_COMMIT = 2;
    int streak = 0;
    // host state at the start of a learned step, restored with the basis when
    // the step is rejected, so the solve continues exactly as it would have
    // without the step
    struct {
        double prev32, vmax;
        int stall32, last_refresh;
        bool phase32, fresh64;
    } snap{};
    if (learned) {
        mlp_begin(c, mlp, nlv);
        d.SnapB.alloc((size_t)nd * nlv);
        if (fast) d.SnapHB.alloc((size_t)nd * nlv);
    }
    PCI_HIP(hipMemsetAsync(d.fail.p, 0, 4 * sizeof(int), s));
    PCI_HIP(hipMemsetAsync(d.conv.p, 0, 64 * sizeof(int), s));
    PCI_HIP(hipMemsetAsync(d.vm.p, 0, 4 * sizeof(double), s));  // vmax = 0 before the first FormP

    // ---- Init4 + DiagInitApprox + FormB0
    struct InitEvents {  // destroyed on every exit path
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~InitEvents() {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
        }
    } tev;
    PCI_HIP(hipEventCreate(&tev.e0));
    PCI_HIP(hipEventCreate(&tev.e1));
    hipEvent_t ti0 = tev.e0, ti1 = tev.e1;
    PCI_HIP(hipEventRecord(ti0, s));
    PCI_HIP(hipMemsetAsync(d.Z1.p, 0, (size_t)nd0 * nd0 * sizeof(double), s));
    head_block_and_diag(c, nd0, d.Z1.p, d.D.p, s);
    if (jsel) k_avg_diag<<<c.nc, 256, 0, s>>>(c.conf_start.p, d.D.p);  // AvgDiag at iteration 1
    {
        DBuf<int> info;
        info.pool = &c.pool;
        info.alloc(1);
        rocblas_status st_;
        // in a batch only the part that needs the whole GPU (the cooperative
        // tridiagonalisation, or FP64 dsyevd) runs one block at a time
        GateScope gate(c, s);
        Eigh32Work ew;
        DBuf<float> e32;
        e32.pool = &c.pool;
        if (fast) {
            // the head-block eigenvectors only shape the preconditioner and
            // the start vectors: FP32 is enough, and faster
            DBuf<float> &z32 = d.Z32;
            z32.alloc((size_t)nd0 * nd0);
            e32.alloc(nd0);
            k_cvt_to32<<<blocks_for((int64_t)nd0 * nd0), 256, 0, s>>>(nd0, d.Z1.p, nd0, z32.p, nd0, nd0);
            // cooperative tridiagonalisation (ssyevd where the block does not
            // fit the GPU's LDS), then sstedc + blocked Q below
            d.eigh_path = eigh32_reduce(c, d.blas, nd0, z32.p, e32.p, ew);
            PCI_HIP(hipMemsetAsync(info.p, 0, sizeof(int), s));
        } else {
            st_ = rocsolver_dsyevd(d.blas, rocblas_evect_original, rocblas_fill_upper, nd0, d.Z1.p, nd0, d.E1.p,
                                   d.S.p /* >= nd0 scratch */, info.p);
            PCI_REQUIRE(st_ == rocblas_status_success, PCI_E_HIP, "rocsolver_dsyevd failed");
            d.Z1T.alloc((size_t)nd0 * nd0);
            transpose_sq<double>(nd0, d.Z1.p, d.Z1T.p, s);
        }
        int hinfo = 0;
        PCI_HIP(hipMemcpyAsync(&hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost, s));
        gate.leave(s);
        if (fast) {
            DBuf<float> &z32 = d.Z32;
            eigh32_finish(c, d.blas, nd0, z32.p, e32.p, ew);
            k_cvt_to64<<<blocks_for((int64_t)nd0 * nd0), 256, 0, s>>>(nd0, z32.p, nd0, d.Z1.p, nd0, nd0);
            k_cvt_to64<<<blocks_for(nd0), 256, 0, s>>>(nd0, e32.p, 1, d.E1.p, 1, 1);
            d.Z32T.alloc((size_t)nd0 * nd0);
            transpose_sq<float>(nd0, z32.p, d.Z32T.p, s);
        }
        PCI_HIP(hipStreamSynchronize(s));  // before the FP32 buffers are released
        PCI_REQUIRE(hinfo == 0, PCI_E_HIP, "dsyevd did not converge");
    }
    std::vector<int> sel;
    DBuf<int> dsel;
    if (jsel) {
        // FormB0 with K_prj = 1: the first Nlv eigenvectors of the Nd0 block
        // whose J (from J^2 on the block) is within 0.1 of XJ_av
        DBuf<double> xj;
        xj.alloc(64);
        std::vector<double> hx(64);
        for (int j0 = 0; j0 < nd0 && (int)sel.size() < nlv; j0 += 64) {
            const int nb = std::min(64, nd0 - j0);
            k_jav_head<<<nb, 256, 0, s>>>(nd0, c.J.rp.p, c.J.col.p, c.J.val.p, d.Z1.p, j0, xj.p);
            PCI_HIP(hipMemcpyAsync(hx.data(), xj.p, nb * sizeof(double), hipMemcpyDeviceToHost, s));
            PCI_HIP(hipStreamSynchronize(s));
            for (int b = 0; b < nb && (int)sel.size() < nlv; b++)
                if (std::fabs(hx[b] - c.xj_av) <= 0.1) sel.push_back(j0 + b);
        }
        PCI_REQUIRE((int)sel.size() == nlv, PCI_E_ARG,
                    "only " + std::to_string(sel.size()) + " eigenvectors of the Nd0 block have J = " +
                        std::to_string(c.xj_av) + ": lower Nlv or raise MaxNd0");
        upload(dsel, sel.data(), nlv, s);
    }
    k_init_B<<<blocks_for((int64_t)nd * d.ldv), 256, 0, s>>>(nd, nd0, nlv, d.ldv, d.Z1.p, jsel ? dsel.p : nullptr,
                                                             d.V.p);
    const bool restart = rs && rs->x0;
    if (restart) {
        // |Kl4| = 2: the start vectors are the caller's (read from CONF.XIJ,
        // src/davidson.f90:136-154) instead of the head eigenvectors; Init4
        // still runs, its eigenpairs shape the correction
        DBuf<double> x0;
        x0.pool = &c.pool;
        upload(x0, rs->x0, (size_t)nlv * nd, s);
        k_set_B<<<blocks_for((int64_t)nd * nlv), 256, 0, s>>>(nd, nlv, d.ldv, x0.p, d.V.p);
        PCI_HIP(hipStreamSynchronize(s));
    }
    PCI_HIP(hipEventRecord(ti1, s));
    PCI_HIP(hipStreamSynchronize(s));
    float tinit = 0.f;
    PCI_HIP(hipEventElapsedTime(&tinit, ti0, ti1));
    std::vector<double> E1h(nd0);
    PCI_HIP(hipMemcpy(E1h.data(), d.E1.p, nd0 * sizeof(double), hipMemcpyDeviceToHost));
    if (jsel)
        for (int j = 0; j < nlv; j++) E1h[j] = E1h[sel[j]];
    std::vector<double> Tk(nlv), Eh(2 * nlv);
    for (int j = 0; j < nlv; j++) { Tk[j] = -(E1h[j] + H.hmin); Eh[j] = E1h[j]; }
    if (restart && rs->tk0)
        for (int j = 0; j < nlv; j++) { Tk[j] = rs->tk0[j]; Eh[j] = -(Tk[j] + H.hmin); }
    // checkpoint: the current vectors (B), Tk and J_av to the caller
    const bool have_j = c.J.valid && c.J.nd == nd && !H.csf;
    auto checkpoint = [&](int it) {
        if (!rs || !rs->cb || rs->every <= 0) return;
        DBuf<double> vo;
        vo.pool = &c.pool;
        vo.alloc((size_t)nlv * nd);
        k_vec_major<<<blocks_for((int64_t)nd * nlv), 256, 0, s>>>(nd, nlv, d.V.p, d.ldv, vo.p);
        std::vector<double> hv((size_t)nlv * nd), tjv(nlv, 0.0), tjo(nlv, 0.0);
        if (have_j) {
            DBuf<double> acc;
            acc.pool = &c.pool;
            acc.alloc(32);
            const int nbj = (nd + 255) / 256;
            k_jav<<<nbj, 256, 0, s>>>(nd, c.J.rp.p, c.J.col.p, c.J.val.p, d.V.p, d.ldv, nlv, d.part.p);
            k_reduce_partials<<<nlv, 256, 0, s>>>(nbj, nlv, d.part.p, acc.p);
            PCI_HIP(hipGetLastError());
            PCI_HIP(hipMemcpyAsync(tjv.data(), acc.p, nlv * sizeof(double), hipMemcpyDeviceToHost, s));
        }
        PCI_HIP(hipMemcpyAsync(hv.data(), vo.p, hv.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        PCI_HIP(hipStreamSynchronize(s));
        for (int j = 0; j < nlv; j++) tjo[j] = have_j ? 0.5 * (std::sqrt(1.0 + tjv[j]) - 1.0) : 0.0;
        rs->cb(rs->user, it, nlv, nd, Tk.data(), tjo.data(), hv.data());
    };
    checkpoint(0);  // FormB0 writes the start vectors (src/davidson.f90:165-179)

    // ---- iterations
    double ax = 1.0, vmax = 0.0, cnx = 1.0;
    // Without the admixture test (PCI_DVD_NOADMIX) P is decomposed on the
    // device (k_jacobi_dev): the host no longer waits for P; it reads the
    // eigenvalues (Tk) and the orthogonalisation flag at its next wait.
    const bool dev_jac = !admix_test;
    bool tk_pending = false;
    auto take_pending = [&]() {  // after a stream synchronisation
        if (!tk_pending) return;
        for (int j = 0; j < nlv; j++) Tk[j] = -(pin_E[j] + H.hmin);
        tk_pending = false;
        PCI_REQUIRE(*pin_fail == 0, PCI_E_ORTHO, "Fail of orthogonalization of the Davidson subspace");
    };
    const double crit = 1e-6;
    const int n2 = 2 * nlv;
    std::vector<double> h_norm(nlv), h_theta(nlv);
    std::vector<int> h_conv(nlv, 0);
    bool phase32 = fast;                   // FP32 SpMM phase of the fast mode
    bool hb_valid = false;                 // fast mode: HV[:, 0:nlv] holds H*B
    // FP32 phase: an FP32 product of eigenvector-like blocks carries an error
    // of ~1e-7 relative to ||H||, so the FP32 residual is trusted down to a
    // few 1e-6; the switch also happens when the FP32 residual stops falling
    const double switch_tol = std::max(30.0 * crt4, 3e-6);
    double prev32 = -1.0;
    int stall32 = 0;
    int it_done = 0, converged = 0, lrn_used = 0, lrn_rej = 0;
    double final_res = 0.0;
    double *Bc = d.V.p, *Cc = d.V.p + nlv, *Tc = d.V.p + 2 * nlv;  // column blocks of V
    double *P11 = d.small.p + 2 * 64 * 64, *P2 = d.small.p + 3 * 64 * 64;
    int last_refresh = 0;
    bool admixture_stop = false;
    bool last_res_fp64 = false;
    for (int it = 1; it <= prm.n_it; it++) {
        double cnx_seen = 0.0;  // learned mode: the residual the step reached (traced)
        if (admix_test && !(ax > crit)) {  // the reference skips all remaining iterations
            admixture_stop = true;
            break;
        }
        it_done = it;
        // Ortn(i), i = 1..Nlv
        d.cholqr(0, nlv, nullptr, (fast && hb_valid) ? 0 : -1, it == 1 ? 2 : 1);
        // Mxmpy(1)
        double *HB;
        int ldhb;
        bool fresh64 = false;  // H*B computed by an FP64 SpMM in this iteration
        if (fast) {
            bool refresh = !hb_valid;
            if (prm.refine_every > 0 && it - last_refresh >= prm.refine_every) refresh = true;
            if (refresh) {
                d.spmm(phase32, Bc, d.ldv, d.HV.p, n2);
                hb_valid = true;
                fresh64 = !phase32;
                last_refresh = it;
            }
            HB = d.HV.p;
            ldhb = n2;
        } else {
            d.spmm(false, Bc, d.ldv, Cc, d.ldv);
            HB = Cc;
            ldhb = d.ldv;
            fresh64 = true;
        }
        // FormP(1), Dvdsn residuals
        // theta, residuals, their norms and the convergence flags on the
        // device; one wait for the host's copy of P11 and the norms
        auto residuals = [&]() {
            d.tn(Bc, d.ldv, nlv, HB, ldhb, nlv, P11);
            k_theta_conv<<<1, 64, 0, s>>>(nlv, P11, nullptr, crt4, d.theta.p, d.conv.p);
            k_residual<<<blocks_for((int64_t)nd * nlv), 256, 0, s>>>(nd, nlv, HB, ldhb, Bc, d.ldv, d.theta.p, Cc,
                                                                     d.ldv);
            d.tn(Cc, d.ldv, nlv, Cc, d.ldv, nlv, d.norms.p);
            k_theta_conv<<<1, 64, 0, s>>>(nlv, P11, d.norms.p, crt4, d.theta.p, d.conv.p);
            PCI_HIP(hipMemcpyAsync(pin_P11, P11, nlv * nlv * sizeof(double), hipMemcpyDeviceToHost, s));
            PCI_HIP(hipMemcpyAsync(pin_G, d.norms.p, nlv * nlv * sizeof(double), hipMemcpyDeviceToHost, s));
            PCI_HIP(hipStreamSynchronize(s));
            take_pending();
            cnx = 0.0;
            for (int j = 0; j < nlv; j++) {
                h_theta[j] = pin_P11[j * nlv + j];
                h_norm[j] = std::sqrt(std::fabs(pin_G[j * nlv + j]));
                cnx = std::max(cnx, h_norm[j]);
            }
        };
        residuals();
        if (fast && phase32) {
            stall32 = (prev32 > 0.0 && cnx > 0.8 * prev32 && cnx < 1e-4) ? stall32 + 1 : 0;
            prev32 = cnx;
        }
        if (fast && phase32 && (cnx <= switch_tol || stall32 >= 2)) {
            // refinement: recompute H*B in FP64 and continue in FP64
            phase32 = false;
            d.spmm(false, Bc, d.ldv, d.HV.p, n2);
            fresh64 = true;
            last_refresh = it;
            residuals();
        }
        if (fast && cnx <= crt4 && !fresh64) {
            // convergence is only declared on an FP64 product
            phase32 = false;
            d.spmm(false, Bc, d.ldv, d.HV.p, n2);
            fresh64 = true;
            last_refresh = it;
            residuals();
        }
        last_res_fp64 = fresh64;
        if (learned) {
            int accepted = 0;
            const bool rej = mlp_observe(mlp, cnx, &accepted, &lrn_rej);
            cnx_seen = cnx;  // the step's own residual, before any rollback
            if (!rej && accepted && snap_valid && ++streak >= LRN_COMMIT) {
                lrn_used += streak;  // committed
                streak = 0;
                snap_valid = false;
            }
            if (rej && snap_valid) {
                // undo the provisional learned steps: back to the basis (and,
                // in fast mode, H*B) the first of them started from and to the
                // host state of that point; the solve then repeats that step
                // with the diagonal correction, exactly as the run without
                // the learned steps
                k_copy_cols<<<blocks_for((int64_t)nd * nlv), 256, 0, s>>>(nd, d.SnapB.p, nlv, Bc, d.ldv, nlv);
                if (fast) {
                    k_copy_cols<<<blocks_for((int64_t)nd * nlv), 256, 0, s>>>(nd, d.SnapHB.p, nlv, d.HV.p, n2, nlv);
                    hb_valid = true;
                } else {
                    d.spmm(false, Bc, d.ldv, Cc, d.ldv);
                }
                prev32 = snap.prev32;
                stall32 = snap.stall32;
                vmax = snap.vmax;
                if (dev_jac)
                    PCI_HIP(hipMemcpyAsync(d.vm.p, d.vm.p + 1, sizeof(double), hipMemcpyDeviceToDevice, s));
                phase32 = snap.phase32;
                last_refresh = snap.last_refresh;
                fresh64 = snap.fresh64;
                residuals();
                last_res_fp64 = fresh64;
                snap_valid = false;
                lrn_rej += streak;  // the provisional steps undone with it
                streak = 0;
            }
        }
        for (int j = 0; j < nlv; j++) h_conv[j] = (h_norm[j] < crt4) ? 1 : 0;
        if (trace && it <= 512 && st) st->resid_hist[it - 1] = learned ? cnx_seen : cnx;
        final_res = cnx;  // (the device flags d.conv were set by residuals())
        // correction vectors (Dvdsn): Z1 projector on rows < nd0 ...
        // 0-based first Z1 column used: n01 = Nlv+1, and 1 when pconf's
        // method starts from given vectors (|Kl4| = 2, src/davidson.f90:
        // 518-519).  The FP32, coupled and learned variants keep Nlv+1 on a
        // restart: with the leading columns, whose eigenvalues lie close to
        // theta, their solves from slightly perturbed vectors took ten times
        // the iterations of a fresh start
        const int n01 = (restart && !fast && !coupled && !learned) ? 0 : nlv;
        const int nz = nd0 - n01;
        if (nz > 0) {
            // S = Z1(:, n01:)^T R(1:nd0, :), scaled by 1/(theta - E1), then
            // B1 = Z1(:, n01:) S  (davidson.f90:450-480)
            // S (row-major nz x nlv) in d.S, then B1 (column-major nd0 x nlv)
            // partial sums: one slice of PJ_ROWS rows per block row, NVB <= 24 wide
            const size_t pneed = (size_t)((nd0 + PJ_ROWS - 1) / PJ_ROWS) * nd0 * 24;
            if (d.projpart.n < pneed) d.projpart.alloc(pneed);
            if (fast) {
                projt_pass<float>(nd0, nz, nlv, d.Z32T.p, nd0, n01, Cc, d.ldv, 1, d.projpart.p, d.S.p, nlv, 1,
                                  d.theta.p, d.E1.p, n01, s);
                projt_pass<float>(nz, nd0, nlv, d.Z32.p + (size_t)n01 * nd0, nd0, 0, d.S.p, nlv, 1, d.projpart.p,
                                  d.B1.p, 1, nd0, nullptr, nullptr, 0, s);
            } else {
                projt_pass<double>(nd0, nz, nlv, d.Z1T.p, nd0, n01, Cc, d.ldv, 1, d.projpart.p, d.S.p, nlv, 1,
                                   d.theta.p, d.E1.p, n01, s);
                projt_pass<double>(nz, nd0, nlv, d.Z1.p + (size_t)n01 * nd0, nd0, 0, d.S.p, nlv, 1, d.projpart.p,
                                   d.B1.p, 1, nd0, nullptr, nullptr, 0, s);
            }
            PCI_HIP(hipGetLastError());
        } else {
            PCI_HIP(hipMemsetAsync(d.B1.p, 0, (size_t)nd0 * nlv * sizeof(double), s));
        }
        // head-to-tail coupling: H[tail, head] applied to the head correction
        // (block Gauss-Seidel on the (head, tail) split of theta - H)
        const double *cpl = nullptr;
        if (coupled && nd0 < nd) {
            k_head_rows<<<blocks_for((int64_t)nd0 * nlv), 256, 0, s>>>(nd0, nlv, d.B1.p, d.Tb.p, d.ldt);
            if (fast) {
                k_cvt_to32<<<blocks_for((int64_t)nd0 * nlv), 256, 0, s>>>(nd0, d.Tb.p, d.ldt, d.Tb32.p, d.ldt32, nlv);
                head_couple_f32(c, nd0, d.Tb32.p, d.ldt32, d.Cp32.p, nlv, nlv, s);
                k_cvt_to64<<<blocks_for((int64_t)nd * nlv), 256, 0, s>>>(nd, d.Cp32.p, nlv, d.Cp.p, nlv, nlv);
            } else {
                head_couple_f64(c, nd0, d.Tb.p, d.ldt, d.Cp.p, nlv, nlv, s);
            }
            cpl = d.Cp.p;
        }
        if (col) mlp_collect_step(c, *col, d.D.p, d.theta.p, Cc, d.ldv, cpl, nlv, Bc, d.ldv, nd0, nlv, h_norm);
        // ... and the diagonal (or learned) correction on the other rows
        const float *lrn = nullptr;
        if (learned) {
            lrn = mlp_apply(c, mlp, d.D.p, d.theta.p, Cc, d.ldv, cpl, nlv, nd0, h_norm);
            if (lrn && !snap_valid) {  // keep the state the learned steps start from
                k_copy_cols<<<blocks_for((int64_t)nd * nlv), 256, 0, s>>>(nd, Bc, d.ldv, d.SnapB.p, nlv, nlv);
                if (fast)
                    k_copy_cols<<<blocks_for((int64_t)nd * nlv), 256, 0, s>>>(nd, d.HV.p, n2, d.SnapHB.p, nlv, nlv);
                snap = {prev32, vmax, stall32, last_refresh, phase32, fresh64};
                if (dev_jac)
                    PCI_HIP(hipMemcpyAsync(d.vm.p + 1, d.vm.p, sizeof(double), hipMemcpyDeviceToDevice, s));
                snap_valid = true;
            }
        }
        k_correction<<<blocks_for((int64_t)nd * nlv), 256, 0, s>>>(nd, nd0, nlv, Cc, d.ldv, d.B1.p, d.D.p,
                                                                   d.theta.p, d.conv.p, lrn, mlp.ldo, cpl, nlv, Cc,
                                                                   d.ldv);
        // normalise, then orthogonalise against B and among themselves
        d.tn(Cc, d.ldv, nlv, Cc, d.ldv, nlv, d.norms.p);
        k_scale_gram<<<blocks_for((int64_t)nd * nlv), 256, 0, s>>>(nd, nlv, Cc, d.ldv, d.norms.p);
        for (int pass = 0; pass < 2; pass++) {
            d.tn(Bc, d.ldv, nlv, Cc, d.ldv, nlv, d.small.p);
            d.nn(Bc, d.ldv, nlv, d.small.p, nlv, Cc, d.ldv, -1.0, 1.0);
        }
        d.cholqr(nlv, nlv, d.conv.p, -1);
        if (jsel) {
            // Prj_J (davidson.f90:743-884): filter the corrections to J = XJ_av
            // with (I - J^2 / (jx (jx + 2))) for the other J terms, largest
            // first, until every unconverged vector has J within 1e-5; then
            // orthogonalise them again (pconf.F90:2109-2128)
            const int jav = (int)(2.0 * c.xj_av + 0.1);
            const int njd = (int)c.jt.size();
            k_copy_cols<<<blocks_for((int64_t)nd * nlv), 256, 0, s>>>(nd, Cc, d.ldv, d.PJx.p, d.ld64, nlv);
            int jx = -1;
            double f = 0.0;
            for (int itr = 1; itr <= njd + 1; itr++) {
                const int itj = njd + 1 - itr;
                if (itj != 0) {
                    jx = c.jt[itj - 1];
                    f = jx > 0 ? 1.0 / ((double)jx * (jx + 2)) : 0.0;
                }
                if (!(jx != jav || itj == 0)) continue;
                spmm_matrix_f64(c.Jf, d.PJx.p, d.ld64, d.PJy.p, d.ld64, nlv, 0.0, s, true);
                d.tn(d.PJx.p, d.ld64, nlv, d.PJy.p, d.ld64, nlv, d.norms.p);
                std::vector<double> g(nlv * nlv);
                PCI_HIP(hipMemcpyAsync(g.data(), d.norms.p, nlv * nlv * sizeof(double), hipMemcpyDeviceToHost, s));
                PCI_HIP(hipStreamSynchronize(s));
                double err1 = 0.0;
                for (int j = 0; j < nlv; j++)
                    if (!h_conv[j]) err1 = std::max(err1, std::fabs(std::sqrt(1.0 + g[j * nlv + j]) - 1.0 - jav));
                if (err1 < 1e-5) break;
                if (itj == 0) break;
                k_jfilter<<<blocks_for((int64_t)nd * nlv), 256, 0, s>>>(nd, nlv, d.PJx.p, d.ld64, d.PJy.p, d.ld64, f,
                                                                        jx == 0 ? 1 : 0, d.conv.p);
                d.tn(d.PJx.p, d.ld64, nlv, d.PJx.p, d.ld64, nlv, d.norms.p);
                PCI_HIP(hipMemcpyAsync(g.data(), d.norms.p, nlv * nlv * sizeof(double), hipMemcpyDeviceToHost, s));
                PCI_HIP(hipStreamSynchronize(s));
                std::vector<double> sc(nlv);
                for (int j = 0; j < nlv; j++) sc[j] = (!h_conv[j] && g[j * nlv + j] > 0) ? 1.0 / std::sqrt(g[j * nlv + j]) : 1.0;
                PCI_HIP(hipMemcpyAsync(d.theta.p + 32, sc.data(), nlv * sizeof(double), hipMemcpyHostToDevice, s));
                k_scale_cols<<<blocks_for((int64_t)nd * nlv), 256, 0, s>>>(nd, nlv, d.PJx.p, d.ld64, d.theta.p + 32);
            }
            k_copy_cols<<<blocks_for((int64_t)nd * nlv), 256, 0, s>>>(nd, d.PJx.p, d.ld64, Cc, d.ldv, nlv);
            for (int pass = 0; pass < 2; pass++) {
                d.tn(Bc, d.ldv, nlv, Cc, d.ldv, nlv, d.small.p);
                d.nn(Bc, d.ldv, nlv, d.small.p, nlv, Cc, d.ldv, -1.0, 1.0);
            }
            d.cholqr(nlv, nlv, d.conv.p, -1);
        }
        // the orthogonalisation failure flag is read with P below (one wait
        // per iteration), or here when the loop ends
        PCI_HIP(hipMemcpyAsync(pin_fail, d.fail.p, sizeof(int), hipMemcpyDeviceToHost, s));
        if (!(cnx > crt4)) {
            PCI_HIP(hipStreamSynchronize(s));
            PCI_REQUIRE(*pin_fail == 0, PCI_E_ORTHO, "Fail of orthogonalization of the Davidson subspace");
            converged = 1;
            if (it == 1) {
                // the reference keeps FormB0's energies when converging at once
            }
            break;
        }
        // Mxmpy(2) + FormP(2)
        double *HC;
        int ldhc;
        if (fast) {
            // H C in FP32 in both phases: after the switch H B is exact to
            // FP64 and the corrections enter the next H B = [HB HC] Q only
            // through coefficients of the order of the residual, so the FP32
            // error of H C stays far below Crt4 (convergence is still
            // confirmed on an FP64 product of B)
            d.spmm(true, Cc, d.ldv, d.HV.p + nlv, n2);
            HC = d.HV.p + nlv;
            ldhc = n2;
        } else {
            d.spmm(false, Cc, d.ldv, Tc, d.ldv);
            HC = Tc;
            ldhc = d.ldv;
        }
        double pmax = h_theta[0];
        for (int j = 1; j < nlv; j++) pmax = std::max(pmax, h_theta[j]);
        vmax = std::max(vmax, pmax);
        d.tn(Bc, d.ldv, n2, HC, ldhc, nlv, P2);
        if (dev_jac) {
            k_assemble_P_dev<<<1, 256, 0, s>>>(nlv, P11, P2, d.conv.p, d.theta.p, d.vm.p, d.P.p);
            k_jacobi_dev<<<1, 256, 0, s>>>(n2, nlv, d.P.p, d.small.p, d.eigs.p, d.vm.p);
            PCI_HIP(hipGetLastError());
            PCI_HIP(hipMemcpyAsync(pin_E, d.eigs.p, nlv * sizeof(double), hipMemcpyDeviceToHost, s));
            tk_pending = true;
            d.nn(Bc, d.ldv, n2, d.small.p, nlv, Tc, d.ldv, 1.0, 0.0);
            k_copy_cols<<<blocks_for((int64_t)nd * nlv), 256, 0, s>>>(nd, Tc, d.ldv, Bc, d.ldv, nlv);
            if (fast) {
                d.nn(d.HV.p, n2, n2, d.small.p, nlv, Tc, d.ldv, 1.0, 0.0);
                k_copy_cols<<<blocks_for((int64_t)nd * nlv), 256, 0, s>>>(nd, Tc, d.ldv, d.HV.p, n2, nlv);
            }
            if (rs && rs->every > 0 && it % rs->every == 0) {
                PCI_HIP(hipStreamSynchronize(s));
                take_pending();
                checkpoint(it);
            }
            continue;
        }
        k_assemble_P<<<1, 256, 0, s>>>(nlv, P11, P2, d.conv.p, vmax, d.P.p);
        PCI_HIP(hipMemcpyAsync(pin_P, d.P.p, n2 * n2 * sizeof(double), hipMemcpyDeviceToHost, s));
        PCI_HIP(hipStreamSynchronize(s));
        PCI_REQUIRE(*pin_fail == 0, PCI_E_ORTHO, "Fail of orthogonalization of the Davidson subspace");
        std::vector<double> Ph(pin_P, pin_P + n2 * n2), Qh;
        jacobi_host(n2, Ph, Eh, Qh);
        ax = 0.0;
        vmax = -1e10;
        for (int i = 0; i < nlv; i++) {
            double xx = 0.0;
            for (int k = 0; k < nlv; k++) xx = std::max(xx, std::fabs(Qh[(k + nlv) * n2 + i]));
            ax = std::max(ax, xx);
            vmax = std::max(vmax, Eh[i]);
        }
        // FormB: B <- [B C] Q(:, 1:Nlv)  (Q row-major n2 x n2: first nlv columns)
        // (pin_Q is rewritten only after the next iteration's wait, when this
        // copy has long completed)
        for (int i = 0; i < n2; i++)
            for (int j = 0; j < nlv; j++) pin_Q[i * nlv + j] = Qh[i * n2 + j];
        PCI_HIP(hipMemcpyAsync(d.small.p, pin_Q, (size_t)n2 * nlv * sizeof(double), hipMemcpyHostToDevice, s));
        d.nn(Bc, d.ldv, n2, d.small.p, nlv, Tc, d.ldv, 1.0, 0.0);
        k_copy_cols<<<blocks_for((int64_t)nd * nlv), 256, 0, s>>>(nd, Tc, d.ldv, Bc, d.ldv, nlv);
        if (fast) {
            d.nn(d.HV.p, n2, n2, d.small.p, nlv, Tc, d.ldv, 1.0, 0.0);
            k_copy_cols<<<blocks_for((int64_t)nd * nlv), 256, 0, s>>>(nd, Tc, d.ldv, d.HV.p, n2, nlv);
        }
        for (int j = 0; j < nlv; j++) Tk[j] = -(Eh[j] + H.hmin);
        // KXIJ: FormB writes CONF.XIJ every KXIJ iterations (src/pconf.F90:2186-2211)
        if (rs && rs->every > 0 && it % rs->every == 0) checkpoint(it);
    }
    if (fast && admixture_stop && !converged) {
        // stopped on the admixture test: report an FP64 residual of the
        // returned vectors (orthonormalised as the next iteration would)
        d.cholqr(0, nlv, nullptr, -1);
        d.spmm(false, Bc, d.ldv, d.HV.p, n2);
        double *HB = d.HV.p;
        const int ldhb = n2;
        d.tn(Bc, d.ldv, nlv, HB, ldhb, nlv, P11);
        std::vector<double> p11(nlv * nlv), g(nlv * nlv), th(nlv);
        PCI_HIP(hipMemcpyAsync(p11.data(), P11, nlv * nlv * sizeof(double), hipMemcpyDeviceToHost, s));
        PCI_HIP(hipStreamSynchronize(s));
        for (int j = 0; j < nlv; j++) th[j] = p11[j * nlv + j];
        PCI_HIP(hipMemcpyAsync(d.theta.p, th.data(), nlv * sizeof(double), hipMemcpyHostToDevice, s));
        k_residual<<<blocks_for((int64_t)nd * nlv), 256, 0, s>>>(nd, nlv, HB, ldhb, Bc, d.ldv, d.theta.p, Cc, d.ldv);
        d.tn(Cc, d.ldv, nlv, Cc, d.ldv, nlv, d.norms.p);
        PCI_HIP(hipMemcpyAsync(g.data(), d.norms.p, nlv * nlv * sizeof(double), hipMemcpyDeviceToHost, s));
        PCI_HIP(hipStreamSynchronize(s));
        final_res = 0.0;
        for (int j = 0; j < nlv; j++) final_res = std::max(final_res, std::sqrt(std::fabs(g[j * nlv + j])));
        last_res_fp64 = true;
    }
    if (!converged && final_res <= crt4 && last_res_fp64) converged = 1;
    if (!converged && admixture_stop) converged = 2;
    // outputs
    // eigenvectors in the caller's vector-major layout, gathered on the device
    DBuf<double> vout;
    vout.pool = &c.pool;
    if (vecs) {
        vout.alloc((size_t)nlv * nd);
        k_vec_major<<<blocks_for((int64_t)nd * nlv), 256, 0, s>>>(nd, nlv, d.V.p, d.ldv, vout.p);
        PCI_HIP(hipMemcpyAsync(vecs, vout.p, (size_t)nlv * nd * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    std::vector<double> tjv(nlv, 0.0);
    // J_av needs J^2 of the basis H is in (none for H in the CSF basis)
    if (have_j) {
        DBuf<double> acc;
        acc.pool = &c.pool;
        acc.alloc(32);
        const int nbj = (nd + 255) / 256;
        k_jav<<<nbj, 256, 0, s>>>(nd, c.J.rp.p, c.J.col.p, c.J.val.p, d.V.p, d.ldv, nlv, d.part.p);
        k_reduce_partials<<<nlv, 256, 0, s>>>(nbj, nlv, d.part.p, acc.p);
        PCI_HIP(hipGetLastError());
        PCI_HIP(hipMemcpyAsync(tjv.data(), acc.p, nlv * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    PCI_HIP(hipStreamSynchronize(s));
    take_pending();  // the last device eigen-decomposition of P, if any
    for (int j = 0; j < nlv; j++) {
        if (tk) tk[j] = Tk[j];
        if (tj) tj[j] = have_j ? 0.5 * (std::sqrt(1.0 + tjv[j]) - 1.0) : 0.0;
    }
    if (learned) mlp_end(c, mlp);
    d.collect_spmm_time();
    if (st) {
        st->iterations = it_done;
        st->converged = converged;
        st->spmm_fp32 = d.n32;
        st->spmm_fp64 = d.n64;
        st->learned_used = lrn_used + streak;  // (provisional steps the solve ended with count)
        st->learned_rejected = lrn_rej;
        st->max_residual = final_res;
        st->t_total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
        st->t_spmm_ms = d.t_spmm;
        st->t_spmm32_ms = d.t_spmm32;
        st->t_init_ms = tinit;
        st->spmm_bytes = d.bytes;
    }
}

}  // namespace pci

using namespace pci;

// Training pairs for the learned correction: run the Davidson procedure
// (diagonal correction) against reference eigenvectors xref (nlv x nd,
// vector-major, e.g. a converged run) and record, for every stride-th
// determinant >= Nd0 and every root, the network features, the residual
// component and the target c_j e_ij, where e_j is the error of the current
// vector (aligned to xref_j, orthogonal to the vector) and
// c_j = <t_diag, e_j> / <e_j, e_j> scales it to the diagonal correction.
extern "C" int pci_gpu_mlp_collect(pci_ctx *ctx, const pci_dvd_params *p, const double *xref, int32_t stride,
                                   int64_t cap, float *x, int32_t *bucket, float *r, float *target, int64_t *n) {
    try {
        PCI_REQUIRE(ctx && p && xref && x && bucket && r && target && n && cap > 0 && stride > 0, PCI_E_ARG,
                    "bad argument");
        Ctx &c = ctx->c;
        PCI_HIP(hipSetDevice(c.device));
        PCI_REQUIRE(c.H.valid, PCI_E_ARG, "build H first");
        const int nd = c.H.nd, nlv = p->nlv;
        std::vector<double> xt((size_t)nd * nlv);
        for (int j = 0; j < nlv; j++)
            for (int64_t i = 0; i < nd; i++) xt[(size_t)i * nlv + j] = xref[(size_t)j * nd + i];
        DBuf<double> dx;
        upload(dx, xt.data(), xt.size(), c.stream);
        MlpCollect col;
        col.xref = dx.p;
        col.cap = cap;
        col.stride = stride;
        col.x.alloc(cap * MLP_NF);
        col.bucket.alloc(cap);
        col.r.alloc(cap);
        col.e.alloc(cap);
        pci_dvd_params q = *p;
        q.flags &= ~(PCI_DVD_LEARNED | PCI_DVD_FP32);  // the base correction: diagonal, or coupled if asked
        std::vector<double> tk(nlv), tj(nlv);
        davidson(c, q, tk.data(), tj.data(), nullptr, nullptr, &col);
        const int64_t m = col.n;
        PCI_HIP(hipMemcpy(x, col.x.p, m * MLP_NF * sizeof(float), hipMemcpyDeviceToHost));
        PCI_HIP(hipMemcpy(bucket, col.bucket.p, m * sizeof(int), hipMemcpyDeviceToHost));
        PCI_HIP(hipMemcpy(r, col.r.p, m * sizeof(float), hipMemcpyDeviceToHost));
        PCI_HIP(hipMemcpy(target, col.e.p, m * sizeof(float), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < m; i++) target[i] *= col.scale[i];
        *n = m;
        return PCI_OK;
    } catch (const pci::Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return PCI_E_HIP;
    }
}

extern "C" int pci_gpu_davidson(pci_ctx *ctx, const pci_dvd_params *p, double *tk, double *tj, double *vecs,
                                pci_dvd_stats *stats) {
    try {
        PCI_REQUIRE(ctx && p, PCI_E_ARG, "null argument");
        PCI_HIP(hipSetDevice(ctx->c.device));
        davidson(ctx->c, *p, tk, tj, vecs, stats);
        return PCI_OK;
    } catch (const pci::Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return PCI_E_HIP;
    }
}

extern "C" int pci_gpu_davidson_ex(pci_ctx *ctx, const pci_dvd_params *p, const double *tk0, const double *x0,
                                   int32_t every, pci_dvd_checkpoint_fn cb, void *user, double *tk, double *tj,
                                   double *vecs, pci_dvd_stats *stats) {
    try {
        PCI_REQUIRE(ctx && p, PCI_E_ARG, "null argument");
        PCI_REQUIRE(every >= 0 && (every == 0 || cb), PCI_E_ARG, "a checkpoint period needs a callback");
        PCI_HIP(hipSetDevice(ctx->c.device));
        DvdRestart rs;
        rs.tk0 = tk0;
        rs.x0 = x0;
        rs.every = every;
        rs.cb = cb;
        rs.user = user;
        davidson(ctx->c, *p, tk, tj, vecs, stats, nullptr, &rs);
        return PCI_OK;
    } catch (const pci::Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return PCI_E_HIP;
    }
}

// Independent CI blocks (e.g. the J / parity blocks of one atom) solved at
// once: one host thread per context, each context's kernels on its own
// stream, so the blocks share the GPU.  The head eigensolves run one block
// at a time on an otherwise idle GPU (BatchGate), the iterations of all
// blocks overlap.
extern "C" int pci_gpu_davidson_batch(int32_t nblocks, pci_ctx *const *ctx, const pci_dvd_params *p,
                                      double *const *tk, double *const *tj, double *const *vecs,
                                      pci_dvd_stats *stats) {
    try {
        PCI_REQUIRE(nblocks >= 0 && (nblocks == 0 || (ctx && p)), PCI_E_ARG, "null argument");
        for (int b = 0; b < nblocks; b++) PCI_REQUIRE(ctx[b], PCI_E_ARG, "null context in the batch");
        for (int b = 0; b < nblocks; b++)
            for (int b2 = 0; b2 < b; b2++)
                PCI_REQUIRE(ctx[b] != ctx[b2], PCI_E_ARG, "a context appears twice in the batch");
        for (int b = 0; b < nblocks; b++)
            PCI_REQUIRE(ctx[b]->c.device == ctx[0]->c.device, PCI_E_ARG, "the blocks of a batch share one device");
        std::vector<int> rc(nblocks, PCI_OK);
        std::vector<std::string> msg(nblocks);
        BatchGate gate;
        gate.n = nblocks;
        for (int b = 0; b < nblocks; b++) {
            ctx[b]->c.gate = &gate;
            ctx[b]->c.gate_arrived = ctx[b]->c.gate_finished = false;
        }
        std::vector<std::thread> th;
        th.reserve(nblocks);
        auto join_all = [&]() {
            for (auto &t : th) t.join();
            for (int b = 0; b < nblocks; b++) ctx[b]->c.gate = nullptr;
        };
        for (int b = 0; b < nblocks; b++) {
            auto body = [&, b]() {
                Ctx &c = ctx[b]->c;
                struct Leave {  // a block that stops early releases the others
                    Ctx &c;
                    ~Leave() {
                        c.gate->arrive(c, false);
                        c.gate->finish(c, false);
                    }
                } leave{c};
                try {
                    PCI_HIP(hipSetDevice(c.device));
                    davidson(c, p[b], tk ? tk[b] : nullptr, tj ? tj[b] : nullptr, vecs ? vecs[b] : nullptr,
                             stats ? stats + b : nullptr);
                } catch (const pci::Error &e) {
                    rc[b] = e.code;
                    msg[b] = e.what();
                } catch (const std::exception &e) {
                    rc[b] = PCI_E_HIP;
                    msg[b] = e.what();
                }
            };
            try {
                th.emplace_back(body);
            } catch (...) {  // no thread for blocks b..: release the started ones
                for (int b2 = b; b2 < nblocks; b2++) {
                    gate.arrive(ctx[b2]->c, false);
                    gate.finish(ctx[b2]->c, false);
                }
                join_all();
                throw;
            }
        }
        join_all();
        for (int b = 0; b < nblocks; b++)
            if (rc[b] != PCI_OK) {
                set_last_error("block " + std::to_string(b) + ": " + msg[b]);
                return rc[b];
            }
        return PCI_OK;
    } catch (const pci::Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return PCI_E_HIP;
    }
}
