// Learned Davidson correction: BF16 MLP on MFMA (gfx950), training and
// inference kernels, the Davidson hooks and the C ABI for training.
//
// Matrix products use v_mfma_f32_32x32x8_bf16: a wavefront computes a 32x32
// FP32 block C += A(32xK) B(Kx32) with A, B in LDS as BF16.  Lane l holds
// A[l%32][k0 + 4(l/32) .. +3] and B[k0 + 4(l/32) .. +3][l%32]; accumulator
// element i of lane l is C[8(i/4) + 4(l/32) + i%4][l%32].
//
// BF16 rounding points (reproduced by the CPU oracle for the gradient
// check): the network input x, the hidden activation h1, the weights W1, W2,
// E, and the back-propagated deltas of both hidden layers before they enter
// an MFMA; h2 and every accumulation stay FP32.
#include <cmath>
#include <cstring>
#include <random>

#include "mlp.h"

namespace pci {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf/nan
    const uint32_t r = ((u >> 16) & 1u) + 0x7fffu;                     // round to nearest even
    return (uint16_t)((u + r) >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) {
    const uint32_t u = (uint32_t)b << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// C(32x32) += A(32xK) * B(Kx32), A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn]
__device__ inline void mfma_block(const uint16_t *A, int sam, int sak, const uint16_t *B, int sbk, int sbn, int K,
                                  f32x16 &acc) {
    const int lane = threadIdx.x & 63, r = lane & 31, h = lane >> 5;
    for (int k0 = 0; k0 < K; k0 += 8) {
        bf16x4 a, b;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int k = k0 + 4 * h + i;
            a[i] = (short)A[r * sam + k * sak];
            b[i] = (short)B[k * sbk + r * sbn];
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
}
__device__ inline int acc_row(int i) { return 8 * (i >> 2) + 4 * ((threadIdx.x & 63) >> 5) + (i & 3); }
__device__ inline int acc_col() { return threadIdx.x & 31; }

// ---------------------------------------------------------------- features
__device__ inline void make_features(double u, double r, double cp, double rms, int rank, float *x) {
    const double s = u >= 0 ? 1.0 : -1.0;
    x[0] = (float)u;
    x[1] = (float)(1.0 / (u + s * 0.02));
    x[2] = (float)tanh(u / 0.1);
    x[3] = (float)log(fabs(u) + 1e-3);
    x[4] = (float)(rank / 3.0);
    const double rr = rms > 0 ? r / rms : 0.0;
    x[5] = (float)rr;
    x[6] = (float)s;
    x[7] = (float)(rms > 0 ? cp / rms : 0.0);
}

// Forward pass of 32 samples held by one wavefront.  sx: 32 x MLP_IN bf16
// input rows (already built), sh1: 32 x MLP_H bf16 scratch.  Returns y for
// sample (lane & 31) in lanes 0..31 (lanes 32..63 return the same values).
// When sh2 != nullptr the FP32 h2 (32 x MLP_H) is stored for backward.
__device__ float mlp_forward32(const uint16_t *sW1, const uint16_t *sW2, const float *b1, const float *b2,
                               const float *w3, float b3, const uint16_t *sx, uint16_t *sh1, float *sh2) {
    const int lane = threadIdx.x & 63;
    // layer 1: h1 = relu(x W1^T + b1)
    for (int nb = 0; nb < MLP_H / 32; nb++) {
        f32x16 acc = {};
        mfma_block(sx, MLP_IN, 1, sW1 + nb * 32 * MLP_IN, 1, MLP_IN, MLP_IN, acc);
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int m = acc_row(i), n = nb * 32 + acc_col();
            const float v = acc[i] + b1[n];
            sh1[m * MLP_H + n] = f2bf(v > 0.f ? v : 0.f);
        }
    }
    __syncthreads();  // every wave of the block calls this function
    // layer 2 and the output dot product
    float ypart[16];
#pragma unroll
    for (int i = 0; i < 16; i++) ypart[i] = 0.f;
    for (int nb = 0; nb < MLP_H / 32; nb++) {
        f32x16 acc = {};
        mfma_block(sh1, MLP_H, 1, sW2 + nb * 32 * MLP_H, 1, MLP_H, MLP_H, acc);
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int m = acc_row(i), n = nb * 32 + acc_col();
            float v = acc[i] + b2[n];
            v = v > 0.f ? v : 0.f;
            if (sh2) sh2[m * MLP_H + n] = v;
            ypart[i] += v * w3[n];
        }
    }
    // sum over the 32 columns held by the lanes of each half-wave
#pragma unroll
    for (int i = 0; i < 16; i++)
        for (int o = 16; o > 0; o >>= 1) ypart[i] += __shfl_xor(ypart[i], o);
    // every lane of half-wave h now holds the sums of rows 8(i/4) + 4h + i%4;
    // fetch the row (lane & 31) from the half that holds it
    const int want = lane & 31;
    const int hsrc = (want >> 2) & 1;
    const int isel = ((want >> 3) << 2) | (want & 3);
    float y = 0.f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const float v = __shfl(ypart[i], hsrc * 32);
        y = (i == isel) ? v : y;
    }
    return y + b3;
}

// ---------------------------------------------------------------- inference
// The forward pass of mlp_forward32 with every operand contiguous in k and
// LDS rows padded by 4 values (the 32 row lanes of an MFMA operand then fall
// on different banks): each operand fragment is one 8-byte LDS read.  Same
// products and roundings as mlp_forward32.
constexpr int MLP_INP = MLP_IN + 4, MLP_HP = MLP_H + 4;  // padded LDS row lengths

__device__ inline void mfma_block_k1(const uint16_t *A, int sam, const uint16_t *B, int sbn, int K, f32x16 &acc) {
    const int lane = threadIdx.x & 63, r = lane & 31, h = lane >> 5;
    for (int k0 = 0; k0 < K; k0 += 8) {
        const bf16x4 a = *(const bf16x4 *)(A + r * sam + k0 + 4 * h);
        const bf16x4 b = *(const bf16x4 *)(B + r * sbn + k0 + 4 * h);
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
}

__device__ float mlp_forward32_padded(const uint16_t *sW1, const uint16_t *sW2, const float *b1, const float *b2,
                                      const float *w3, float b3, const uint16_t *sx, uint16_t *sh1) {
    const int lane = threadIdx.x & 63;
    for (int nb = 0; nb < MLP_H / 32; nb++) {
        f32x16 acc = {};
        mfma_block_k1(sx, MLP_INP, sW1 + nb * 32 * MLP_INP, MLP_INP, MLP_IN, acc);
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int m = acc_row(i), n = nb * 32 + acc_col();
            const float v = acc[i] + b1[n];
            sh1[m * MLP_HP + n] = f2bf(v > 0.f ? v : 0.f);
        }
    }
    __syncthreads();  // every wave of the block calls this function
    float ypart[16];
#pragma unroll
    for (int i = 0; i < 16; i++) ypart[i] = 0.f;
    for (int nb = 0; nb < MLP_H / 32; nb++) {
        f32x16 acc = {};
        mfma_block_k1(sh1, MLP_HP, sW2 + nb * 32 * MLP_HP, MLP_HP, MLP_H, acc);
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int n = nb * 32 + acc_col();
            float v = acc[i] + b2[n];
            v = v > 0.f ? v : 0.f;
            ypart[i] += v * w3[n];
        }
    }
#pragma unroll
    for (int i = 0; i < 16; i++)
        for (int o = 16; o > 0; o >>= 1) ypart[i] += __shfl_xor(ypart[i], o);
    const int want = lane & 31;
    const int hsrc = (want >> 2) & 1;
    const int isel = ((want >> 3) << 2) | (want & 3);
    float y = 0.f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const float v = __shfl(ypart[i], hsrc * 32);
        y = (i == isel) ? v : y;
    }
    return y + b3;
}

// Four wavefronts per block, 32 samples per wavefront; sample s -> (row
// nd0 + s / nlv, root s % nlv).  The block keeps the BF16 weights in LDS and
// walks groups of 4 tiles with a grid stride, so the weights are read once
// per block rather than once per tile.
__global__ __launch_bounds__(256) void k_mlp_infer(int nd, int nd0, int nlv, const double *__restrict__ D,
                                                   const double *__restrict__ theta, const double *__restrict__ R,
                                                   int ldr, const double *__restrict__ Cp, int ldc,
                                                   const float *__restrict__ rms,
                                                   const int *__restrict__ rank, const int *__restrict__ bucket,
                                                   const float *__restrict__ w, const uint16_t *__restrict__ wb,
                                                   float *__restrict__ out, int ldo) {
    __shared__ __align__(16) uint16_t sW1[MLP_H * MLP_INP];
    __shared__ __align__(16) uint16_t sW2[MLP_H * MLP_HP];
    __shared__ __align__(16) uint16_t sx[4][32 * MLP_INP];
    __shared__ __align__(16) uint16_t sh1[4][32 * MLP_HP];
    for (int e = threadIdx.x; e < MLP_H * MLP_IN; e += blockDim.x) sW1[(e / MLP_IN) * MLP_INP + e % MLP_IN] = wb[e];
    for (int e = threadIdx.x; e < MLP_H * MLP_H; e += blockDim.x)
        sW2[(e / MLP_H) * MLP_HP + e % MLP_H] = wb[MLP_H * MLP_IN + e];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t nsamp = (int64_t)(nd - nd0) * nlv;
    const int64_t ngroups = (nsamp + 127) / 128;
    const uint16_t *E = wb + MLP_H * MLP_IN + MLP_H * MLP_H;
    uint16_t *x = sx[wv];
    for (int64_t grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {  // uniform over the block
        const int64_t s0 = (grp * 4 + wv) * 32;
        float r_s = 0.f;
        int64_t i_s = -1;
        int j_s = 0;
        if (lane < 32) {
            const int64_t s = s0 + lane;
            float f[MLP_NF];
            for (int q = 0; q < MLP_NF; q++) f[q] = 0.f;
            int bk = 0;
            if (s < nsamp) {
                i_s = nd0 + s / nlv;
                j_s = (int)(s % nlv);
                const double r = R[i_s * ldr + j_s];
                const double cp = Cp ? Cp[i_s * ldc + j_s] : 0.0;
                r_s = (float)(r + cp);
                make_features(D[i_s] - theta[j_s], r, cp, rms[j_s], rank[i_s], f);
                bk = bucket[i_s];
            }
            for (int q = 0; q < MLP_NF; q++) x[lane * MLP_INP + q] = f2bf(f[q]);
            for (int q = 0; q < MLP_NE; q++) x[lane * MLP_INP + MLP_NF + q] = E[bk * MLP_NE + q];
            for (int q = MLP_NF + MLP_NE; q < MLP_IN; q++) x[lane * MLP_INP + q] = 0;
        }
        __syncthreads();
        const float y = mlp_forward32_padded(sW1, sW2, w + MLP_OFF_B1, w + MLP_OFF_B2, w + MLP_OFF_W3, w[MLP_OFF_B3],
                                             x, sh1[wv]);
        if (lane < 32 && i_s >= 0) {
            const float f1 = bf2f(x[lane * MLP_INP + 1]);
            out[i_s * ldo + j_s] = r_s * (-f1 + y);
        }
        __syncthreads();  // x and h1 are rewritten by the next group
    }
}

// Forward over explicit features (x: n x MLP_NF FP32), t = r g per sample.
__global__ __launch_bounds__(256) void k_mlp_fwd(int64_t n, const float *__restrict__ X, const int *__restrict__ bucket,
                                                 const float *__restrict__ Rr, const float *__restrict__ w,
                                                 const uint16_t *__restrict__ wb, float *__restrict__ out) {
    __shared__ __align__(16) uint16_t sW1[MLP_H * MLP_INP];
    __shared__ __align__(16) uint16_t sW2[MLP_H * MLP_HP];
    __shared__ __align__(16) uint16_t sx[4][32 * MLP_INP];
    __shared__ __align__(16) uint16_t sh1[4][32 * MLP_HP];
    for (int e = threadIdx.x; e < MLP_H * MLP_IN; e += blockDim.x) sW1[(e / MLP_IN) * MLP_INP + e % MLP_IN] = wb[e];
    for (int e = threadIdx.x; e < MLP_H * MLP_H; e += blockDim.x)
        sW2[(e / MLP_H) * MLP_HP + e % MLP_H] = wb[MLP_H * MLP_IN + e];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t s0 = ((int64_t)blockIdx.x * 4 + wv) * 32;
    const uint16_t *E = wb + MLP_H * MLP_IN + MLP_H * MLP_H;
    uint16_t *x = sx[wv];
    if (lane < 32) {
        const int64_t s = s0 + lane;
        const bool ok = s < n;
        const int bk = ok ? bucket[s] : 0;
        for (int q = 0; q < MLP_NF; q++) x[lane * MLP_INP + q] = f2bf(ok ? X[s * MLP_NF + q] : 0.f);
        for (int q = 0; q < MLP_NE; q++) x[lane * MLP_INP + MLP_NF + q] = E[bk * MLP_NE + q];
        for (int q = MLP_NF + MLP_NE; q < MLP_IN; q++) x[lane * MLP_INP + q] = 0;
    }
    __syncthreads();
    // the forward pass of the Davidson inference kernel (k_mlp_infer), so the
    // oracle comparison of this entry point covers it
    const float y = mlp_forward32_padded(sW1, sW2, w + MLP_OFF_B1, w + MLP_OFF_B2, w + MLP_OFF_W3, w[MLP_OFF_B3], x,
                                         sh1[wv]);
    if (lane < 32 && s0 + lane < n) out[s0 + lane] = Rr[s0 + lane] * (-bf2f(x[lane * MLP_INP + 1]) + y);
}

// ---------------------------------------------------------------- training
// Forward + backward over one tile of 32 samples per block (one wavefront);
// gradients accumulated with atomics into g.  loss += inv_n * sum (r g - target)^2
// and g += gradient of inv_b * sum (r g - target)^2 (inv_b: 1 / batch size).
__global__ __launch_bounds__(64) void k_mlp_train(int64_t n, const float *__restrict__ X,
                                                  const int *__restrict__ bucket, const float *__restrict__ Rr,
                                                  const float *__restrict__ T, const float *__restrict__ w,
                                                  const uint16_t *__restrict__ wb, float *__restrict__ g,
                                                  float *__restrict__ loss, float inv_n, float inv_b) {
    __shared__ uint16_t sW1[MLP_H * MLP_IN];
    __shared__ uint16_t sW2[MLP_H * MLP_H];
    __shared__ uint16_t sx[32 * MLP_IN];
    __shared__ uint16_t sh1[32 * MLP_H];
    __shared__ float sh2[32 * MLP_H];
    __shared__ uint16_t sd2[32 * MLP_H];   // delta of layer 2 (pre-activation), bf16
    __shared__ uint16_t sd1[32 * MLP_H];   // delta of layer 1 (pre-activation), bf16
    __shared__ float sdy[32];
    for (int e = threadIdx.x; e < MLP_H * MLP_IN; e += 64) sW1[e] = wb[e];
    for (int e = threadIdx.x; e < MLP_H * MLP_H; e += 64) sW2[e] = wb[MLP_H * MLP_IN + e];
    const int lane = threadIdx.x & 63;
    const int64_t s0 = (int64_t)blockIdx.x * 32;
    const uint16_t *E = wb + MLP_H * MLP_IN + MLP_H * MLP_H;
    int bk = 0;
    bool valid = false;
    if (lane < 32) {
        const int64_t s = s0 + lane;
        valid = s < n;
        bk = valid ? bucket[s] : 0;
        for (int q = 0; q < MLP_NF; q++) sx[lane * MLP_IN + q] = f2bf(valid ? X[s * MLP_NF + q] : 0.f);
        for (int q = 0; q < MLP_NE; q++) sx[lane * MLP_IN + MLP_NF + q] = valid ? E[bk * MLP_NE + q] : 0;
        for (int q = MLP_NF + MLP_NE; q < MLP_IN; q++) sx[lane * MLP_IN + q] = 0;
    }
    __syncthreads();
    const float y = mlp_forward32(sW1, sW2, w + MLP_OFF_B1, w + MLP_OFF_B2, w + MLP_OFF_W3, w[MLP_OFF_B3], sx, sh1,
                                  sh2);
    __syncthreads();
    // output layer: t = r (-f1 + y); dL/dy = 2 (t - target) r / n
    if (lane < 32) {
        float dy = 0.f, l = 0.f;
        if (valid) {
            const int64_t s = s0 + lane;
            const float r = Rr[s];
            const float t = r * (-bf2f(sx[lane * MLP_IN + 1]) + y);
            const float d = t - T[s];
            l = d * d * inv_n;
            dy = 2.f * d * r * inv_b;
        }
        sdy[lane] = dy;
        for (int o = 16; o > 0; o >>= 1) l += __shfl_xor(l, o, 32);
        if (lane == 0) atomicAdd(loss, l);
    }
    __syncthreads();
    // dw3, db3, delta2 = dy w3 [h2 > 0]
    for (int c = lane; c < MLP_H; c += 64) {
        float acc = 0.f;
        for (int m = 0; m < 32; m++) acc += sdy[m] * sh2[m * MLP_H + c];
        atomicAdd(&g[MLP_OFF_W3 + c], acc);
    }
    if (lane == 0) {
        float acc = 0.f;
        for (int m = 0; m < 32; m++) acc += sdy[m];
        atomicAdd(&g[MLP_OFF_B3], acc);
    }
    const float *w3 = w + MLP_OFF_W3;
    for (int e = lane; e < 32 * MLP_H; e += 64) {
        const int m = e / MLP_H, c = e % MLP_H;
        const float d = sh2[e] > 0.f ? sdy[m] * w3[c] : 0.f;
        sd2[e] = f2bf(d);
    }
    __syncthreads();
    // db2, dW2 = delta2^T h1 (128 x 128, K = 32 samples)
    for (int c = lane; c < MLP_H; c += 64) {
        float acc = 0.f;
        for (int m = 0; m < 32; m++) acc += bf2f(sd2[m * MLP_H + c]);
        atomicAdd(&g[MLP_OFF_B2 + c], acc);
    }
    for (int ob = 0; ob < MLP_H / 32; ob++)
        for (int ib = 0; ib < MLP_H / 32; ib++) {
            f32x16 acc = {};
            mfma_block(sd2 + ob * 32, 1, MLP_H, sh1 + ib * 32, MLP_H, 1, 32, acc);
#pragma unroll
            for (int i = 0; i < 16; i++)
                atomicAdd(&g[MLP_OFF_W2 + (int64_t)(ob * 32 + acc_row(i)) * MLP_H + ib * 32 + acc_col()], acc[i]);
        }
    // delta1 = (delta2 W2) [h1 > 0]
    for (int nb = 0; nb < MLP_H / 32; nb++) {
        f32x16 acc = {};
        mfma_block(sd2, MLP_H, 1, sW2 + nb * 32, MLP_H, 1, MLP_H, acc);
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int m = acc_row(i), c = nb * 32 + acc_col();
            sd1[m * MLP_H + c] = f2bf(bf2f(sh1[m * MLP_H + c]) > 0.f ? acc[i] : 0.f);
        }
    }
    __syncthreads();
    // db1, dW1 = delta1^T x (128 x 32, K = 32 samples)
    for (int c = lane; c < MLP_H; c += 64) {
        float acc = 0.f;
        for (int m = 0; m < 32; m++) acc += bf2f(sd1[m * MLP_H + c]);
        atomicAdd(&g[MLP_OFF_B1 + c], acc);
    }
    for (int ob = 0; ob < MLP_H / 32; ob++) {
        f32x16 acc = {};
        mfma_block(sd1 + ob * 32, 1, MLP_H, sx, MLP_IN, 1, 32, acc);
#pragma unroll
        for (int i = 0; i < 16; i++)
            atomicAdd(&g[MLP_OFF_W1 + (int64_t)(ob * 32 + acc_row(i)) * MLP_IN + acc_col()], acc[i]);
    }
    // dE[bucket] += (delta1 W1)[:, 8:24]
    {
        f32x16 acc = {};
        mfma_block(sd1, MLP_H, 1, sW1, MLP_IN, 1, MLP_H, acc);
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int m = acc_row(i), c = acc_col();
            const int64_t s = s0 + m;
            if (c >= MLP_NF && c < MLP_NF + MLP_NE && s < n)
                atomicAdd(&g[MLP_OFF_E + (int64_t)bucket[s] * MLP_NE + (c - MLP_NF)], acc[i]);
        }
    }
}

__global__ void k_adam(int64_t np, float *__restrict__ w, float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1t, float b2t) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi;
    const float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / b1t) / (sqrtf(vi / b2t) + 1e-8f);
    g[i] = 0.f;
}

__global__ void k_to_bf16(const float *__restrict__ w, uint16_t *__restrict__ wb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n1 = MLP_H * MLP_IN, n2 = MLP_H * MLP_H, ne = MLP_NB * MLP_NE;
    if (i < n1) wb[i] = f2bf(w[MLP_OFF_W1 + i]);
    else if (i < n1 + n2) wb[i] = f2bf(w[MLP_OFF_W2 + i - n1]);
    else if (i < n1 + n2 + ne) wb[i] = f2bf(w[MLP_OFF_E + i - n1 - n2]);
}

// ---------------------------------------------------------------- collection
// Aligned error of the current vectors against the reference eigenvectors:
// e_j = s_j xref_j - b_j, s_j = sign(<xref_j, b_j>), then e_j -= <e_j,b_j> b_j.
__global__ void k_err(int nd, int nlv, const double *__restrict__ xref, const double *__restrict__ B, int ldb,
                      const double *__restrict__ sgn, double *__restrict__ E) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nd * nlv) return;
    const int64_t r = i / nlv;
    const int j = (int)(i % nlv);
    E[i] = sgn[j] * xref[i] - B[r * ldb + j];
}
__global__ void k_err_proj(int nd, int nlv, const double *__restrict__ B, int ldb, const double *__restrict__ eb,
                           double *__restrict__ E, const double *__restrict__ D, const double *__restrict__ theta,
                           const double *__restrict__ R, int ldr, const double *__restrict__ Cp, int ldc, int nd0,
                           double *__restrict__ TD) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nd * nlv) return;
    const int64_t r = i / nlv;
    const int j = (int)(i % nlv);
    const double e = E[i] - eb[j] * B[r * ldb + j];
    const bool tail = r >= nd0;
    E[i] = tail ? e : 0.0;
    TD[i] = tail ? (R[r * ldr + j] + (Cp ? Cp[r * ldc + j] : 0.0)) / (theta[j] - D[r]) : 0.0;
}
__global__ void k_collect(int nd, int nd0, int nlv, int stride, const double *__restrict__ D,
                          const double *__restrict__ theta, const double *__restrict__ R, int ldr,
                          const double *__restrict__ Cp, int ldc,
                          const float *__restrict__ rms, const int *__restrict__ rank, const int *__restrict__ bucket,
                          const double *__restrict__ E, int64_t base, int64_t cap, float *__restrict__ X,
                          int *__restrict__ Bk, float *__restrict__ Rr, float *__restrict__ Ee) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nrow = (nd - nd0 + stride - 1) / stride;
    if (t >= nrow * nlv) return;
    const int64_t i = nd0 + (t / nlv) * stride;
    const int j = (int)(t % nlv);
    const int64_t o = base + t;
    if (o >= cap) return;
    float f[MLP_NF];
    const double r = R[i * ldr + j];
    const double cp = Cp ? Cp[i * ldc + j] : 0.0;
    make_features(D[i] - theta[j], r, cp, rms[j], rank[i], f);
    for (int q = 0; q < MLP_NF; q++) X[o * MLP_NF + q] = f[q];
    Bk[o] = bucket[i];
    Rr[o] = (float)(r + cp);
    Ee[o] = (float)E[i * nlv + j];
}

// ---------------------------------------------------------------- host side
static MlpNet &net(Ctx &c) {
    if (!c.mlp) c.mlp = new MlpNet();
    return *c.mlp;
}

void mlp_free(Ctx &c) {
    delete c.mlp;
    c.mlp = nullptr;
}

static void net_alloc(Ctx &c) {
    MlpNet &N = net(c);
    if (N.w.p) return;
    N.w.alloc(MLP_NPARAM);
    N.wb.alloc(MLP_H * MLP_IN + MLP_H * MLP_H + MLP_NB * MLP_NE);
    N.g.alloc(MLP_NPARAM);
    N.m.alloc(MLP_NPARAM);
    N.v.alloc(MLP_NPARAM);
}

static void net_set(Ctx &c, const float *h) {
    net_alloc(c);
    MlpNet &N = net(c);
    hipStream_t s = c.stream;
    PCI_HIP(hipMemcpyAsync(N.w.p, h, MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice, s));
    PCI_HIP(hipMemsetAsync(N.g.p, 0, MLP_NPARAM * sizeof(float), s));
    PCI_HIP(hipMemsetAsync(N.m.p, 0, MLP_NPARAM * sizeof(float), s));
    PCI_HIP(hipMemsetAsync(N.v.p, 0, MLP_NPARAM * sizeof(float), s));
    const int nb = MLP_H * MLP_IN + MLP_H * MLP_H + MLP_NB * MLP_NE;
    k_to_bf16<<<(nb + 255) / 256, 256, 0, s>>>(N.w.p, N.wb.p);
    PCI_HIP(hipGetLastError());
    PCI_HIP(hipStreamSynchronize(s));
    N.step = 0;
    N.ready = true;
}

void mlp_begin(Ctx &c, MlpPrecond &m, int nlv) {
    PCI_REQUIRE(c.mlp && c.mlp->ready, PCI_E_ARG, "learned correction requested but no network is loaded");
    m.nlv = nlv;
    m.ldo = nlv;
    m.out.pool = m.rms.pool = &c.pool;  // per-solve buffers from the context's pool
    m.out.alloc((size_t)c.nd * nlv);
    m.rms.alloc(64);
    m.disabled = false;
    m.tried = false;
    m.seen = 0;
    m.prev_cnx = -1.0;
    m.rate_diag = -1.0;
}

const float *mlp_apply(Ctx &c, MlpPrecond &m, const double *D, const double *theta, const double *R, int ldr,
                       const double *C, int ldc, int nd0, const std::vector<double> &rnorm) {
    if (m.disabled || nd0 >= c.nd || m.seen < MLP_WARMUP || m.rate_diag < 0.0) {
        m.tried = false;
        return nullptr;
    }
    std::vector<float> rms(m.nlv);
    for (int j = 0; j < m.nlv; j++) rms[j] = (float)(rnorm[j] / std::sqrt((double)c.nd));
    PCI_HIP(hipMemcpyAsync(m.rms.p, rms.data(), m.nlv * sizeof(float), hipMemcpyHostToDevice, c.stream));
    const int64_t ns = (int64_t)(c.nd - nd0) * m.nlv;
    const unsigned grid = (unsigned)std::min<int64_t>((ns + 127) / 128, 1024);
    k_mlp_infer<<<grid, 256, 0, c.stream>>>(c.nd, nd0, m.nlv, D, theta, R, ldr, C, ldc, m.rms.p, c.det_rank.p,
                                            c.det_bucket.p, c.mlp->w.p, c.mlp->wb.p, m.out.p, m.ldo);
    PCI_HIP(hipGetLastError());
    m.tried = true;
    return m.out.p;
}

bool mlp_observe(MlpPrecond &m, double cnx, int *used, int *rejected) {
    bool rej = false;
    if (m.prev_cnx > 0.0) {
        const double rate = cnx / m.prev_cnx;
        if (m.tried) {
            if (rate < 1.0 && rate <= m.rate_diag) (*used)++;
            else {
                (*rejected)++;
                m.disabled = true;  // diagonal correction from here on
                rej = true;
            }
        } else {
            m.rate_diag = rate;
        }
    }
    if (!rej) m.prev_cnx = cnx;  // a rejected step is undone by the caller
    m.seen++;
    return rej;
}

void mlp_end(Ctx &, MlpPrecond &m) {
    m.out.release();
    m.rms.release();
}

void mlp_collect_step(Ctx &c, MlpCollect &col, const double *D, const double *theta, const double *R, int ldr,
                      const double *C, int ldc, const double *B, int ldb, int nd0, int nlv,
                      const std::vector<double> &rnorm) {
    if (col.n >= col.cap || nd0 >= c.nd) return;
    hipStream_t s = c.stream;
    const int nd = c.nd;
    const size_t ne = (size_t)nd * nlv;
    DBuf<double> E, TD, dots;
    E.alloc(ne);
    TD.alloc(ne);
    dots.alloc(4 * 64);
    std::vector<double> xb(nlv), hb(nlv);
    // per-root dot products on the host side of small reductions
    auto dot_cols = [&](const double *A, int lda, const double *Bm, int ldbm, std::vector<double> &out) {
        std::vector<double> ha((size_t)nd * nlv), hbm((size_t)nd * nlv);
        std::vector<double> ta((size_t)nd * lda), tb((size_t)nd * ldbm);
        PCI_HIP(hipMemcpyAsync(ta.data(), A, ta.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        PCI_HIP(hipMemcpyAsync(tb.data(), Bm, tb.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        PCI_HIP(hipStreamSynchronize(s));
        for (int j = 0; j < nlv; j++) {
            double acc = 0.0;
            for (int64_t r = 0; r < nd; r++) acc += ta[r * lda + j] * tb[r * ldbm + j];
            out[j] = acc;
        }
    };
    dot_cols(col.xref, nlv, B, ldb, xb);
    std::vector<double> sg(nlv);
    for (int j = 0; j < nlv; j++) sg[j] = xb[j] >= 0 ? 1.0 : -1.0;
    PCI_HIP(hipMemcpyAsync(dots.p, sg.data(), nlv * sizeof(double), hipMemcpyHostToDevice, s));
    k_err<<<(unsigned)((ne + 255) / 256), 256, 0, s>>>(nd, nlv, col.xref, B, ldb, dots.p, E.p);
    std::vector<double> eb(nlv);
    dot_cols(E.p, nlv, B, ldb, eb);
    PCI_HIP(hipMemcpyAsync(dots.p + 64, eb.data(), nlv * sizeof(double), hipMemcpyHostToDevice, s));
    k_err_proj<<<(unsigned)((ne + 255) / 256), 256, 0, s>>>(nd, nlv, B, ldb, dots.p + 64, E.p, D, theta, R, ldr,
                                                             C, ldc, nd0, TD.p);
    std::vector<double> te(nlv), ee(nlv);
    dot_cols(TD.p, nlv, E.p, nlv, te);
    dot_cols(E.p, nlv, E.p, nlv, ee);
    std::vector<float> rms(nlv);
    for (int j = 0; j < nlv; j++) rms[j] = (float)(rnorm[j] / std::sqrt((double)nd));
    DBuf<float> drms;
    upload(drms, rms.data(), nlv, s);
    const int64_t nrow = (nd - nd0 + col.stride - 1) / col.stride;
    const int64_t cnt = std::min(nrow * nlv, col.cap - col.n);
    k_collect<<<(unsigned)((nrow * nlv + 255) / 256), 256, 0, s>>>(nd, nd0, nlv, col.stride, D, theta, R, ldr,
                                                                   C, ldc, drms.p, c.det_rank.p, c.det_bucket.p, E.p, col.n,
                                                                   col.cap, col.x.p, col.bucket.p, col.r.p, col.e.p);
    PCI_HIP(hipGetLastError());
    PCI_HIP(hipStreamSynchronize(s));
    // target = c_j e with c_j = <t_diag, e_j> / <e_j, e_j>, recorded per sample
    for (int64_t t = 0; t < cnt; t++) {
        const int j = (int)(t % nlv);
        col.scale.push_back(ee[j] > 0 ? (float)(te[j] / ee[j]) : 0.f);
    }
    col.n += cnt;
}

}  // namespace pci

// ---------------------------------------------------------------- C ABI
using namespace pci;

#define API_BEGIN try {
#define API_END                                                   \
    return PCI_OK;                                                \
    }                                                             \
    catch (const pci::Error &e) {                                 \
        set_last_error(e.what());                                 \
        return e.code;                                            \
    }                                                             \
    catch (const std::exception &e) {                             \
        set_last_error(e.what());                                 \
        return PCI_E_HIP;                                         \
    }

extern "C" int pci_gpu_mlp_nparams(int64_t *n) {
    API_BEGIN
    PCI_REQUIRE(n, PCI_E_ARG, "null argument");
    *n = MLP_NPARAM;
    API_END
}

extern "C" int pci_gpu_mlp_init(pci_ctx *ctx, uint64_t seed) {
    API_BEGIN
    PCI_REQUIRE(ctx, PCI_E_ARG, "null context");
    PCI_HIP(hipSetDevice(ctx->c.device));
    std::vector<float> w(MLP_NPARAM, 0.f);
    std::mt19937_64 gen(seed);
    std::normal_distribution<float> nd(0.f, 1.f);
    for (int64_t i = 0; i < MLP_H * MLP_IN; i++)
        w[MLP_OFF_W1 + i] = (i % MLP_IN) < MLP_NF + MLP_NE ? nd(gen) * std::sqrt(2.f / (MLP_NF + MLP_NE)) : 0.f;
    for (int64_t i = 0; i < MLP_H * MLP_H; i++) w[MLP_OFF_W2 + i] = nd(gen) * std::sqrt(2.f / MLP_H);
    for (int64_t i = 0; i < MLP_H; i++) w[MLP_OFF_W3 + i] = nd(gen) * 1e-3f;
    for (int64_t i = 0; i < MLP_NB * MLP_NE; i++) w[MLP_OFF_E + i] = nd(gen) * 0.1f;
    net_set(ctx->c, w.data());
    API_END
}

extern "C" int pci_gpu_mlp_set_weights(pci_ctx *ctx, const float *w, int64_t n) {
    API_BEGIN
    PCI_REQUIRE(ctx && w && n == MLP_NPARAM, PCI_E_ARG, "weights must hold pci_gpu_mlp_nparams values");
    PCI_HIP(hipSetDevice(ctx->c.device));
    net_set(ctx->c, w);
    API_END
}

extern "C" int pci_gpu_mlp_get_weights(pci_ctx *ctx, float *w, int64_t n) {
    API_BEGIN
    PCI_REQUIRE(ctx && w && n == MLP_NPARAM, PCI_E_ARG, "buffer must hold pci_gpu_mlp_nparams values");
    PCI_REQUIRE(ctx->c.mlp && ctx->c.mlp->ready, PCI_E_ARG, "no network");
    PCI_HIP(hipSetDevice(ctx->c.device));
    PCI_HIP(hipMemcpy(w, ctx->c.mlp->w.p, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    API_END
}

// Loss and gradient (or an Adam step) over n samples.
static void run_batch(Ctx &c, int64_t n, const float *dX, const int *dB, const float *dR, const float *dT,
                      float *dloss, float inv_n) {
    MlpNet &N = net(c);
    const unsigned grid = (unsigned)((n + 31) / 32);
    k_mlp_train<<<grid, 64, 0, c.stream>>>(n, dX, dB, dR, dT, N.w.p, N.wb.p, N.g.p, dloss, inv_n, 1.0f / (float)n);
    PCI_HIP(hipGetLastError());
}

extern "C" int pci_gpu_mlp_grad(pci_ctx *ctx, const float *x, const int32_t *bucket, const float *r,
                                const float *target, int64_t n, float *grad, float *loss) {
    API_BEGIN
    PCI_REQUIRE(ctx && x && bucket && r && target && grad && n > 0, PCI_E_ARG, "null argument");
    Ctx &c = ctx->c;
    PCI_HIP(hipSetDevice(c.device));
    PCI_REQUIRE(c.mlp && c.mlp->ready, PCI_E_ARG, "no network");
    for (int64_t i = 0; i < n; i++) PCI_REQUIRE(bucket[i] >= 0 && bucket[i] < MLP_NB, PCI_E_ARG, "bucket out of range");
    hipStream_t s = c.stream;
    DBuf<float> dX, dR, dT, dl;
    DBuf<int> dB;
    upload(dX, x, n * MLP_NF, s);
    upload(dB, bucket, n, s);
    upload(dR, r, n, s);
    upload(dT, target, n, s);
    dl.alloc(1);
    PCI_HIP(hipMemsetAsync(dl.p, 0, sizeof(float), s));
    PCI_HIP(hipMemsetAsync(c.mlp->g.p, 0, MLP_NPARAM * sizeof(float), s));
    run_batch(c, n, dX.p, dB.p, dR.p, dT.p, dl.p, 1.0f / (float)n);
    PCI_HIP(hipMemcpyAsync(grad, c.mlp->g.p, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, s));
    float hl = 0.f;
    PCI_HIP(hipMemcpyAsync(&hl, dl.p, sizeof(float), hipMemcpyDeviceToHost, s));
    PCI_HIP(hipMemsetAsync(c.mlp->g.p, 0, MLP_NPARAM * sizeof(float), s));
    PCI_HIP(hipStreamSynchronize(s));
    if (loss) *loss = hl;
    API_END
}

extern "C" int pci_gpu_mlp_train(pci_ctx *ctx, const float *x, const int32_t *bucket, const float *r,
                                 const float *target, int64_t n, int32_t epochs, int32_t batch, float lr,
                                 float *loss_hist) {
    API_BEGIN
    PCI_REQUIRE(ctx && x && bucket && r && target && n > 0 && epochs > 0 && batch > 0, PCI_E_ARG, "bad argument");
    Ctx &c = ctx->c;
    PCI_HIP(hipSetDevice(c.device));
    PCI_REQUIRE(c.mlp && c.mlp->ready, PCI_E_ARG, "initialise the network first");
    for (int64_t i = 0; i < n; i++) PCI_REQUIRE(bucket[i] >= 0 && bucket[i] < MLP_NB, PCI_E_ARG, "bucket out of range");
    MlpNet &N = *c.mlp;
    hipStream_t s = c.stream;
    DBuf<float> dX, dR, dT, dl;
    DBuf<int> dB;
    upload(dX, x, n * MLP_NF, s);
    upload(dB, bucket, n, s);
    upload(dR, r, n, s);
    upload(dT, target, n, s);
    dl.alloc(epochs);
    PCI_HIP(hipMemsetAsync(dl.p, 0, epochs * sizeof(float), s));
    const int nb_w = MLP_H * MLP_IN + MLP_H * MLP_H + MLP_NB * MLP_NE;
    for (int ep = 0; ep < epochs; ep++) {
        for (int64_t b0 = 0; b0 < n; b0 += batch) {
            const int64_t nb = std::min<int64_t>(batch, n - b0);
            run_batch(c, nb, dX.p + b0 * MLP_NF, dB.p + b0, dR.p + b0, dT.p + b0, dl.p + ep, 1.0f / (float)n);
            N.step++;
            const float b1t = 1.f - std::pow(0.9f, (float)N.step), b2t = 1.f - std::pow(0.999f, (float)N.step);
            k_adam<<<(unsigned)((MLP_NPARAM + 255) / 256), 256, 0, s>>>(MLP_NPARAM, N.w.p, N.g.p, N.m.p, N.v.p, lr,
                                                                        b1t, b2t);
            k_to_bf16<<<(nb_w + 255) / 256, 256, 0, s>>>(N.w.p, N.wb.p);
            PCI_HIP(hipGetLastError());
        }
    }
    if (loss_hist) PCI_HIP(hipMemcpyAsync(loss_hist, dl.p, epochs * sizeof(float), hipMemcpyDeviceToHost, s));
    PCI_HIP(hipStreamSynchronize(s));
    API_END
}

extern "C" int pci_gpu_mlp_forward(pci_ctx *ctx, const float *x, const int32_t *bucket, const float *r, int64_t n,
                                   float *t) {
    API_BEGIN
    PCI_REQUIRE(ctx && x && bucket && r && t && n > 0, PCI_E_ARG, "null argument");
    Ctx &c = ctx->c;
    PCI_HIP(hipSetDevice(c.device));
    PCI_REQUIRE(c.mlp && c.mlp->ready, PCI_E_ARG, "no network");
    for (int64_t i = 0; i < n; i++) PCI_REQUIRE(bucket[i] >= 0 && bucket[i] < MLP_NB, PCI_E_ARG, "bucket out of range");
    hipStream_t s = c.stream;
    DBuf<float> dX, dR, dO;
    DBuf<int> dB;
    upload(dX, x, n * MLP_NF, s);
    upload(dB, bucket, n, s);
    upload(dR, r, n, s);
    dO.alloc(n);
    const int64_t tiles = (n + 31) / 32;
    k_mlp_fwd<<<(unsigned)((tiles + 3) / 4), 256, 0, s>>>(n, dX.p, dB.p, dR.p, c.mlp->w.p, c.mlp->wb.p, dO.p);
    PCI_HIP(hipGetLastError());
    std::vector<float> out(n);
    PCI_HIP(hipMemcpyAsync(out.data(), dO.p, n * sizeof(float), hipMemcpyDeviceToHost, s));
    PCI_HIP(hipStreamSynchronize(s));
    memcpy(t, out.data(), n * sizeof(float));
    API_END
}
