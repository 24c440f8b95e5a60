// Dense symmetric eigensolver for the Nd0 block of the fast mode (Init4 +
// DiagInitApprox, src/davidson.f90:20-73, src/pconf.F90:1866-1891: DSYEVD of
// Z1 = H[1:Nd0, 1:Nd0] - Hmin I, here in FP32).
//
// rocSOLVER's ssyevd spends most of its time in the Householder
// tridiagonalisation (ssytrd), a chain of n small launches per column.  Here
// that step is one cooperative kernel: the matrix is held in LDS across the
// whole GPU (column j on workgroup j mod G, at most TRD_C columns of at most
// TRD_T * TRD_R rows per workgroup), and each column costs one grid-wide
// barrier:
//
//   P1  (own columns j > i) apply the pending rank-2 update of step i-1 and
//       form y_j = A_i(:, j) . v_i (A_i symmetric, so y = A_i v_i by columns);
//       the owner of column i+1 publishes it (updates up to step i-1);
//   --  grid barrier: per-workgroup arrival flags, workgroup 0 releases --
//   P2  (every workgroup, redundantly and identically) w_i = tau y - tau/2
//       (tau y . v) v, column i+1 after update i = published - v w(i+1) - w,
//       and the Householder reflector of column i+1 (LAPACK slarfg);
//       workgroup 0 writes d, e, tau and the reflector in LAPACK's 'L' layout.
//
// The output is what ssytrd('L') returns, so rocSOLVER's sstedc (eigen-
// vectors of the tridiagonal matrix) and apply_q (Q applied in blocks,
// below) finish the solve.
//
// Everything workgroups exchange (y, the published column, the flags) is
// written and read with agent-scope atomic stores and loads, which are
// coherent across the XCDs' L2 caches; a writer waits for its stores to
// complete (s_waitcnt) before its workgroup raises its flag, and readers
// load only after seeing the release.  No cache-wide write-back or
// invalidate (a __threadfence) is needed per barrier: those cost ~10 us
// per barrier with 256 workgroups, the flags ~2.5 us.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "internal.h"

namespace pci {

constexpr int TRD_T = 1024;  // threads per workgroup
constexpr int TRD_R = 3;     // rows per thread: n <= 3072
constexpr int TRD_C = 12;    // columns per workgroup
constexpr int TRD_W = TRD_T / 64;

struct TrdArgs {
    int n, cpw;
    const float *A;     // input, column-major n x n (full symmetric)
    float *V;           // output: reflectors below the first subdiagonal (ld n)
    float *d, *e, *tau;
    float *y, *colpub;  // exchange vectors (2n each: one n-slice per column parity)
    unsigned *flags;    // arrival flag of workgroup b at flags[64 b] (zeroed by the caller)
    unsigned *gen;      // release generation written by workgroup 0 (zeroed)
    int *abort;         // set when a barrier wait exceeds its bound
    unsigned spin_max;  // the bound: sleeps per wait (2^22 unless PCI_TRD_SPIN sets it, for tests)
};

__device__ __forceinline__ void xstore(float *p, float v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float xload(const float *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// wavefront sum by DPP: row shifts 1, 2, 4, 8 (zero fill), then the row
// broadcasts 15 and 31; the total ends in lane 63
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_src(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL, ROW_MASK, 0xf, true));
}
__device__ __forceinline__ float wave_sum63(float x) {
    x += dpp_src<0x111, 0xf>(x);  // row_shr:1
    x += dpp_src<0x112, 0xf>(x);  // row_shr:2
    x += dpp_src<0x114, 0xf>(x);  // row_shr:4
    x += dpp_src<0x118, 0xf>(x);  // row_shr:8
    x += dpp_src<0x142, 0xa>(x);  // row_bcast:15
    x += dpp_src<0x143, 0xc>(x);  // row_bcast:31
    return x;
}

// grid-wide barrier: every workgroup raises its flag once its exchange
// stores have completed; workgroup 0 waits for all flags (one per thread)
// and releases the generation the others wait for.  Waits are bounded: on
// timeout the abort flag is raised and the kernel ends at the next barrier.
__device__ bool trd_sync(const TrdArgs &a, unsigned step) {
    __shared__ int ok;
    const unsigned G = gridDim.x, b = blockIdx.x, t = threadIdx.x;
    if (t == 0) ok = 1;
    __builtin_amdgcn_s_waitcnt(0);  // this thread's exchange stores are complete
    __syncthreads();
    if (t == 0) __hip_atomic_store(a.flags + 64 * b, step + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (b == 0) {
        if (t < G) {
            unsigned spins = 0;
            while (__hip_atomic_load(a.flags + 64 * t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < step + 1) {
                __builtin_amdgcn_s_sleep(1);
                if (++spins > a.spin_max) {
                    __hip_atomic_store(a.abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = 0;
                    break;
                }
            }
        }
        __syncthreads();
        if (t == 0) __hip_atomic_store(a.gen, step + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else if (t == 0) {
        unsigned spins = 0;
        while (__hip_atomic_load(a.gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < step + 1) {
            __builtin_amdgcn_s_sleep(1);
            if (((++spins) & 1023u) == 0 &&
                (spins > a.spin_max || __hip_atomic_load(a.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                __hip_atomic_store(a.abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
        }
    }
    __syncthreads();
    return ok != 0;
}

// sum over the workgroup of one value per thread; every thread gets the sum
// (fixed order: wave shuffles, then the waves in index order).  Each call
// site passes its own red buffer, used once per column, so the barriers
// between two uses of a buffer already guarantee its readers are done.
__device__ float trd_sum(float x, float *red) {
    x = wave_sum63(x);
    if ((threadIdx.x & 63) == 63) red[threadIdx.x >> 6] = x;
    __syncthreads();
    float s = 0.f;
#pragma unroll
    for (int w = 0; w < TRD_W; w++) s += red[w];
    return s;
}

// Householder reflector of column c (rows c+1..n-1 held in col[], one slot
// per thread row), LAPACK slarfg: v(c+1) = 1, v(r) = col(r) / (alpha - beta)
// below, tau = (beta - alpha) / beta, beta = -sign(alpha) |(alpha, x)|.
__device__ void trd_house(int n, int c, const int *rr, const float *col, float *v, float &tau, float &beta,
                          float *red, float *bc) {
    // alpha = col(c+1), broadcast from the thread holding that row
#pragma unroll
    for (int q = 0; q < TRD_R; q++)
        if (rr[q] == c + 1) bc[0] = col[q];
    float xs = 0.f;
#pragma unroll
    for (int q = 0; q < TRD_R; q++)
        if (rr[q] >= c + 2 && rr[q] < n) xs += col[q] * col[q];
    const float xn2 = trd_sum(xs, red);  // (its barriers also publish bc[0])
    const float alpha = bc[0];
    if (xn2 == 0.f) {
        tau = 0.f;
        beta = alpha;
#pragma unroll
        for (int q = 0; q < TRD_R; q++) v[q] = (rr[q] == c + 1) ? 1.f : 0.f;
        return;
    }
    beta = -copysignf(sqrtf(alpha * alpha + xn2), alpha);
    tau = (beta - alpha) / beta;
    const float sc = 1.f / (alpha - beta);
#pragma unroll
    for (int q = 0; q < TRD_R; q++) {
        const int r = rr[q];
        v[q] = (r == c + 1) ? 1.f : ((r >= c + 2 && r < n) ? col[q] * sc : 0.f);
    }
}

__global__ __launch_bounds__(TRD_T) void k_trd(TrdArgs a) {
    extern __shared__ float lds[];  // cpw columns of n floats
    __shared__ float red[TRD_W * TRD_C];
    __shared__ float red_dot[TRD_W], red_house[TRD_W];  // the two block sums of a column
    __shared__ float bc[4];
    __shared__ float vjp[TRD_C], wjp[TRD_C];  // pending v, w at the rows of the own columns
    const int n = a.n, cpw = a.cpw, G = gridDim.x, b = blockIdx.x, t = threadIdx.x;
    const int lane = t & 63, wv = t >> 6;
    int rr[TRD_R];
#pragma unroll
    for (int q = 0; q < TRD_R; q++) rr[q] = t + q * TRD_T;

    // own columns into LDS
    for (int s = 0; s < cpw; s++) {
        const int j = b + s * G;
        if (j >= n) break;
        for (int r = t; r < n; r += TRD_T) lds[(size_t)s * n + r] = a.A[(size_t)j * n + r];
    }
    // step 0: reflector of column 0 (the input), by every workgroup
    float v[TRD_R], vp[TRD_R], wp[TRD_R], col[TRD_R];
#pragma unroll
    for (int q = 0; q < TRD_R; q++) {
        col[q] = (rr[q] >= 1 && rr[q] < n) ? a.A[rr[q]] : 0.f;
        vp[q] = 0.f;
        wp[q] = 0.f;
    }
    float tau, beta;
    trd_house(n, 0, rr, col, v, tau, beta, red_house, bc);
    if (b == 0) {
        if (t == 0) {
            a.d[0] = a.A[0];
            a.e[0] = beta;
            a.tau[0] = tau;
        }
#pragma unroll
        for (int q = 0; q < TRD_R; q++) {
            const int r = rr[q];
            if (r >= 2 && r < n) a.V[r] = v[q];
            if (r == 1) a.V[r] = beta;
        }
    }
    if (t < TRD_C) {
        vjp[t] = 0.f;
        wjp[t] = 0.f;
    }
    __syncthreads();

    for (int i = 0; i < n - 1; i++) {
        // exchange slices of this column.  One barrier separates this
        // column's stores (P1) from its loads (P2); a workgroup that leaves
        // it early may store the next column's values before a slower one
        // has loaded these, so consecutive columns use different slices, and
        // the next barrier orders these loads before column i + 2's stores
        float *const yb = a.y + (size_t)(i & 1) * n;
        float *const cb = a.colpub + (size_t)(i & 1) * n;
        // ---- P1: pending update of step i-1 on the own columns j > i, y_j
        float acc[TRD_C];
#pragma unroll
        for (int s = 0; s < TRD_C; s++) acc[s] = 0.f;
#pragma unroll
        for (int s = 0; s < TRD_C; s++) {
            const int j = b + s * G;
            if (s >= cpw || j >= n || j <= i) continue;
            const float vj = vjp[s], wj = wjp[s];
            float *cs = lds + (size_t)s * n;
#pragma unroll
            for (int q = 0; q < TRD_R; q++) {
                const int r = rr[q];
                if (r > i && r < n) {
                    const float x = cs[r] - (vp[q] * wj + wp[q] * vj);
                    cs[r] = x;
                    acc[s] += x * v[q];
                }
            }
        }
        // publish column i+1 (owner only)
        if ((i + 1) % G == b) {
            const float *cs = lds + (size_t)((i + 1) / G) * n;
#pragma unroll
            for (int q = 0; q < TRD_R; q++)
                if (rr[q] > i && rr[q] < n) xstore(cb + rr[q], cs[rr[q]]);
        }
        // y_j = column sums of acc
#pragma unroll
        for (int s = 0; s < TRD_C; s++) {
            const float x = wave_sum63(acc[s]);
            if (lane == 63) red[wv * TRD_C + s] = x;
        }
        __syncthreads();
        if (t < cpw) {
            const int j = b + t * G;
            if (j > i && j < n) {
                float s = 0.f;
                for (int w = 0; w < TRD_W; w++) s += red[w * TRD_C + t];
                xstore(yb + j, s);
            }
        }
        if (!trd_sync(a, (unsigned)i)) return;

        // ---- P2 (every workgroup): w_i, column i+1 after update i, reflector i+1
        float p[TRD_R], w[TRD_R];
        float ps = 0.f;
#pragma unroll
        for (int q = 0; q < TRD_R; q++) {
            const int r = rr[q];
            const bool in = r > i && r < n;
            p[q] = in ? tau * xload(yb + r) : 0.f;
            col[q] = in ? xload(cb + r) : 0.f;
            ps += p[q] * v[q];
        }
        // p(i+1) rides on the sum's barrier; w(i+1) = p(i+1) + alpha2 (v(i+1) = 1)
#pragma unroll
        for (int q = 0; q < TRD_R; q++)
            if (rr[q] == i + 1) bc[1] = p[q];
        const float sdot = trd_sum(ps, red_dot);
        const float alpha2 = -0.5f * tau * sdot;
#pragma unroll
        for (int q = 0; q < TRD_R; q++) w[q] = p[q] + alpha2 * v[q];
        const float w1 = bc[1] + alpha2;
#pragma unroll
        for (int q = 0; q < TRD_R; q++) {
            const int r = rr[q];
            col[q] = (r > i && r < n) ? col[q] - (v[q] * w1 + w[q]) : 0.f;
        }
        // pending update for the next P1, at the rows of the own columns
#pragma unroll
        for (int q = 0; q < TRD_R; q++) {
            const int r = rr[q];
            vp[q] = v[q];
            wp[q] = w[q];
            if (r < n && r % G == b && r / G < TRD_C) {
                vjp[r / G] = v[q];
                wjp[r / G] = w[q];
            }
        }
        if (i + 1 == n - 1) {  // last diagonal element
            if (b == 0)
#pragma unroll
                for (int q = 0; q < TRD_R; q++)
                    if (rr[q] == n - 1) a.d[n - 1] = col[q];
            break;
        }
        if (b == 0)
#pragma unroll
            for (int q = 0; q < TRD_R; q++)
                if (rr[q] == i + 1) a.d[i + 1] = col[q];
        // reflector of column i+1 from rows i+2.. (col holds rows i+1..)
        float cn[TRD_R];
#pragma unroll
        for (int q = 0; q < TRD_R; q++) cn[q] = (rr[q] >= i + 2) ? col[q] : 0.f;
        float vn[TRD_R];
        trd_house(n, i + 1, rr, cn, vn, tau, beta, red_house, bc);
#pragma unroll
        for (int q = 0; q < TRD_R; q++) v[q] = vn[q];
        if (b == 0) {
            if (t == 0) {
                a.e[i + 1] = beta;
                a.tau[i + 1] = tau;
            }
            float *vc = a.V + (size_t)(i + 1) * n;
#pragma unroll
            for (int q = 0; q < TRD_R; q++) {
                const int r = rr[q];
                if (r >= i + 3 && r < n) vc[r] = v[q];
                if (r == i + 2) vc[r] = beta;
            }
        }
        __syncthreads();  // vjp/wjp written before the next P1 reads them
    }
}

// Back-transformation Z <- Q Z of the tridiagonal eigenvectors, Q = H(0) ..
// H(n-2) from k_trd (ssytrd('L') layout: reflector j in column j of V at
// rows j+2.., its unit at row j+1), in blocks of WY_NB reflectors applied
// last block first: Q_b = I - V_b T_b V_b^T with T_b^{-1} = strict_upper(V_b^T
// V_b) + diag(1 / tau) (the inverse of the forward-columnwise larft
// recurrence), so a block costs three sgemm, one strsm and two small
// kernels.  A reflector with tau = 0 is the identity: its column of V_b is
// zero and its 1/tau is taken as 1.
constexpr int WY_NB = 128;

__global__ void k_wy_v(int n, int i0, int k, int m, const float *__restrict__ V, const float *__restrict__ tau,
                       float *__restrict__ Vb) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m * k) return;
    const int j = (int)(t / m), r = (int)(t - (int64_t)j * m);
    float x = 0.f;
    if (tau[i0 + j] != 0.f) x = r == j ? 1.f : (r > j ? V[(size_t)(i0 + j) * n + (i0 + 1 + r)] : 0.f);
    Vb[t] = x;
}

__global__ void k_wy_tinv(int k, int i0, const float *__restrict__ G, const float *__restrict__ tau,
                          float *__restrict__ Ti) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= k * k) return;
    const int j = t / k, i = t - j * k;  // column-major (row i, column j)
    const float tj = tau[i0 + j];
    Ti[t] = j > i ? G[t] : (j == i ? (tj != 0.f ? 1.f / tj : 1.f) : 0.f);
}

static void apply_q(rocblas_handle h, hipStream_t s, int n, const float *V, const float *tau, float *Z,
                    BlockPool *pool) {
    const int nref = n - 1;
    if (nref <= 0) return;
    DBuf<float> Vb, G, Ti, W;
    Vb.pool = G.pool = Ti.pool = W.pool = pool;
    Vb.alloc((size_t)n * WY_NB);
    G.alloc((size_t)WY_NB * WY_NB);
    Ti.alloc((size_t)WY_NB * WY_NB);
    W.alloc((size_t)WY_NB * n);
    const float one = 1.f, zero = 0.f, mone = -1.f;
    for (int b = (nref + WY_NB - 1) / WY_NB - 1; b >= 0; b--) {
        const int i0 = b * WY_NB, k = std::min(WY_NB, nref - i0), m = n - 1 - i0;
        k_wy_v<<<(unsigned)(((int64_t)m * k + 255) / 256), 256, 0, s>>>(n, i0, k, m, V, tau, Vb.p);
        PCI_REQUIRE(rocblas_sgemm(h, rocblas_operation_transpose, rocblas_operation_none, k, k, m, &one, Vb.p, m,
                                  Vb.p, m, &zero, G.p, k) == rocblas_status_success,
                    PCI_E_HIP, "rocblas_sgemm failed");
        k_wy_tinv<<<(unsigned)((k * k + 255) / 256), 256, 0, s>>>(k, i0, G.p, tau, Ti.p);
        float *Zs = Z + i0 + 1;
        PCI_REQUIRE(rocblas_sgemm(h, rocblas_operation_transpose, rocblas_operation_none, k, n, m, &one, Vb.p, m, Zs,
                                  n, &zero, W.p, k) == rocblas_status_success,
                    PCI_E_HIP, "rocblas_sgemm failed");
        PCI_REQUIRE(rocblas_strsm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                                  rocblas_diagonal_non_unit, k, n, &one, Ti.p, k, W.p, k) == rocblas_status_success,
                    PCI_E_HIP, "rocblas_strsm failed");
        PCI_REQUIRE(rocblas_sgemm(h, rocblas_operation_none, rocblas_operation_none, m, n, k, &mone, Vb.p, m, W.p, k,
                                  &one, Zs, n) == rocblas_status_success,
                    PCI_E_HIP, "rocblas_sgemm failed");
    }
    PCI_HIP(hipGetLastError());
}


// Eigenvalues W (ascending) and eigenvectors (overwriting A, column-major,
// ld n) of a symmetric n x n FP32 matrix on the context stream.  The
// cooperative tridiagonalisation is used when the matrix fits the GPU's LDS
// (n <= 3072 and at most TRD_C columns per compute unit); otherwise, or when
// the cooperative launch is refused, rocSOLVER's ssyevd.  Returns the path
// taken: 1 cooperative, 0 ssyevd.
int eigh32_reduce(Ctx &c, rocblas_handle h, int n, float *A, float *W, Eigh32Work &w) {
    hipStream_t s = c.stream;
    if (c.cu_count == 0) {
        hipDeviceProp_t prop;
        PCI_HIP(hipGetDeviceProperties(&prop, c.device));
        c.coop_launch = prop.cooperativeLaunch != 0;
        c.lds_per_block = prop.sharedMemPerBlock;
        c.cu_count = prop.multiProcessorCount;
    }
    const int G = c.cu_count;
    const int cpw = (n + G - 1) / G;
    const size_t shm = (size_t)cpw * n * sizeof(float);
    bool coop = n >= 3 && n <= TRD_T * TRD_R && cpw <= TRD_C && G <= TRD_T && c.coop_launch &&
                shm + 2048 <= c.lds_per_block && !std::getenv("PCI_EIGH_ROCSOLVER");
    if (coop) {
        if (hipFuncSetAttribute((const void *)k_trd, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm) !=
            hipSuccess)
            coop = false;
        int occ = 0;
        if (coop && (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_trd, TRD_T, shm) != hipSuccess || occ < 1))
            coop = false;
        (void)hipGetLastError();
    }
    w.V.pool = w.e.pool = w.tau.pool = w.y.pool = w.colpub.pool = &c.pool;
    w.flags.pool = &c.pool;
    w.abort_flag.pool = w.info.pool = &c.pool;
    w.info.alloc(1);
    if (coop) {
        w.V.alloc((size_t)n * n);
        w.e.alloc(n);
        w.tau.alloc(n);
        w.y.alloc((size_t)2 * n);
        w.colpub.alloc((size_t)2 * n);
        w.flags.alloc((size_t)64 * (G + 1));
        w.abort_flag.alloc(1);
        PCI_HIP(hipMemsetAsync(w.flags.p, 0, (size_t)64 * (G + 1) * sizeof(unsigned), s));
        PCI_HIP(hipMemsetAsync(w.abort_flag.p, 0, sizeof(int), s));
        PCI_HIP(hipMemsetAsync(w.V.p, 0, (size_t)n * n * sizeof(float), s));
        const char *sp = std::getenv("PCI_TRD_SPIN");
        const unsigned spin_max = sp ? (unsigned)std::strtoul(sp, nullptr, 10) : (1u << 22);
        TrdArgs a{n,      cpw,      A,       w.V.p, W, w.e.p, w.tau.p, w.y.p, w.colpub.p, w.flags.p,
                  w.flags.p + (size_t)64 * G, w.abort_flag.p, spin_max};
        void *args[] = {&a};
        // refused (e.g. fewer co-resident workgroups than compute units): ssyevd
        if (hipLaunchCooperativeKernel((const void *)k_trd, dim3(G), dim3(TRD_T), args, (unsigned)shm, s) !=
            hipSuccess) {
            (void)hipGetLastError();
            coop = false;
        }
    }
    w.coop = coop;
    if (!coop) {
        DBuf<float> e2;
        e2.alloc(n);
        rocblas_status st = rocsolver_ssyevd(h, rocblas_evect_original, rocblas_fill_upper, n, A, n, W, e2.p,
                                             w.info.p);
        PCI_REQUIRE(st == rocblas_status_success, PCI_E_HIP, "rocsolver_ssyevd failed");
        int hinfo = 0;
        PCI_HIP(hipMemcpyAsync(&hinfo, w.info.p, sizeof(int), hipMemcpyDeviceToHost, s));
        PCI_HIP(hipStreamSynchronize(s));
        PCI_REQUIRE(hinfo == 0, PCI_E_HIP, "ssyevd did not converge");
        return 0;
    }
    return 1;
}

void eigh32_finish(Ctx &c, rocblas_handle h, int n, float *A, float *W, Eigh32Work &w) {
    if (!w.coop) return;  // ssyevd already gave the eigenvectors
    hipStream_t s = c.stream;
    DBuf<float> Z;
    Z.pool = &c.pool;
    Z.alloc((size_t)n * n);
    int hab = 0;
    PCI_HIP(hipMemcpyAsync(&hab, w.abort_flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    PCI_HIP(hipStreamSynchronize(s));
    if (hab) {
        // a grid barrier wait ran out (the kernel ended early without a
        // result); k_trd only reads A, so rocSOLVER's ssyevd solves it instead
        w.coop = false;
        DBuf<float> e2;
        e2.pool = &c.pool;
        e2.alloc(n);
        rocblas_status st = rocsolver_ssyevd(h, rocblas_evect_original, rocblas_fill_upper, n, A, n, W, e2.p,
                                             w.info.p);
        PCI_REQUIRE(st == rocblas_status_success, PCI_E_HIP, "rocsolver_ssyevd failed");
        int hinfo = 0;
        PCI_HIP(hipMemcpyAsync(&hinfo, w.info.p, sizeof(int), hipMemcpyDeviceToHost, s));
        PCI_HIP(hipStreamSynchronize(s));
        PCI_REQUIRE(hinfo == 0, PCI_E_HIP, "ssyevd did not converge");
        c.trd_fallbacks++;
        return;
    }
    // eigenvectors of the tridiagonal matrix, then Q applied (ssyevd's last steps)
    rocblas_status st = rocsolver_sstedc(h, rocblas_evect_tridiagonal, n, W, w.e.p, Z.p, n, w.info.p);
    PCI_REQUIRE(st == rocblas_status_success, PCI_E_HIP, "rocsolver_sstedc failed");
    if (std::getenv("PCI_EIGH_SORMTR")) {
        st = rocsolver_sormtr(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, n, n, w.V.p, n,
                              w.tau.p, Z.p, n);
        PCI_REQUIRE(st == rocblas_status_success, PCI_E_HIP, "rocsolver_sormtr failed");
    } else {
        apply_q(h, s, n, w.V.p, w.tau.p, Z.p, &c.pool);
    }
    PCI_HIP(hipMemcpyAsync(A, Z.p, (size_t)n * n * sizeof(float), hipMemcpyDeviceToDevice, s));
    int hinfo = 0;
    PCI_HIP(hipMemcpyAsync(&hinfo, w.info.p, sizeof(int), hipMemcpyDeviceToHost, s));
    PCI_HIP(hipStreamSynchronize(s));
    PCI_REQUIRE(hinfo == 0, PCI_E_HIP, "sstedc did not converge");
}

// Returns the path: 1 cooperative tridiagonalisation, 0 ssyevd, 2 the
// cooperative kernel gave up at a barrier and ssyevd solved it instead.
int dense_eigh32(Ctx &c, rocblas_handle h, int n, float *A, float *W) {
    Eigh32Work w;
    const int path = eigh32_reduce(c, h, n, A, W, w);
    eigh32_finish(c, h, n, A, W, w);
    return path == 1 && !w.coop ? 2 : path;
}

}  // namespace pci

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

extern "C" int pci_gpu_eigh32(pci_ctx *ctx, int32_t n, float *a, float *w, int32_t *path) {
    API_BEGIN
    PCI_REQUIRE(ctx && a && w && n >= 1, PCI_E_ARG, "pci_gpu_eigh32: bad arguments");
    Ctx &c = ctx->c;
    PCI_HIP(hipSetDevice(c.device));
    rocblas_handle h = nullptr;
    PCI_REQUIRE(rocblas_create_handle(&h) == rocblas_status_success, PCI_E_HIP, "rocblas_create_handle failed");
    struct Guard {
        rocblas_handle h;
        ~Guard() { rocblas_destroy_handle(h); }
    } g{h};
    rocblas_set_stream(h, c.stream);
    DBuf<float> A, W;
    upload(A, a, (size_t)n * n, c.stream);
    W.alloc(n);
    const int p = dense_eigh32(c, h, n, A.p, W.p);
    PCI_HIP(hipMemcpyAsync(a, A.p, (size_t)n * n * sizeof(float), hipMemcpyDeviceToHost, c.stream));
    PCI_HIP(hipMemcpyAsync(w, W.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c.stream));
    PCI_HIP(hipStreamSynchronize(c.stream));
    if (path) *path = p;
    API_END
}
