// Block SpMM Y = (H - Hmin I) X over a block of Davidson vectors.
//
// Reference: Mxmpy (src/davidson.f90:572-647) applies the stored lower
// triangle to one column at a time by symmetric scatter.  On the device the
// matrix is held with both triangles (full rows), so every output row is a
// gather over its own row with no atomics, and all nvec columns of the block
// are applied in one pass over the matrix: the matrix stream (col + value,
// 8 or 12 bytes per element) is read once per block instead of once per
// column.  X and Y are row-major with leading dimensions ldx / ldy, so the
// nvec values X[col, :] gathered for one element are contiguous.
//
// Launch shape: one 64-lane wavefront per row (rows of H hold ~1e3 elements
// at the target sizes); lanes stride over the row's elements, each keeping
// nvec partial sums in registers, followed by a butterfly reduction.  Rows
// are assigned so consecutive workgroups stay on one XCD's L2 for the
// X gathers (blockIdx round-robins over the 8 XCDs).
#include "internal.h"

namespace pci {

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Map blockIdx so that the blocks dispatched to one XCD (workgroups go to
// the 8 XCDs round-robin by blockIdx) cover one contiguous range of rows.
// nb must be a multiple of 8, which makes the map a permutation of [0, nb).
__device__ inline int xcd_block(int b, int nb) {
    const int per = nb >> 3;
    return (b & 7) * per + (b >> 3);
}

// Scalar-gather kernel: one lane per element, NV scalar loads of X[col, :].
// Used when X rows are not 16-byte aligned or not padded.
template <typename VT, typename XT, typename AT, int NV>
__global__ __launch_bounds__(256) void k_spmm(int nd, const int64_t *__restrict__ rp, const int *__restrict__ col,
                                              const VT *__restrict__ val, const XT *__restrict__ X, int ldx,
                                              XT *__restrict__ Y, int ldy, AT shift) {
    const int lane = threadIdx.x & 63;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int row = b * 4 + (threadIdx.x >> 6);
    if (row >= nd) return;
    AT acc[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) acc[v] = 0;
    const int64_t j0 = rp[row], j1 = rp[row + 1];
    for (int64_t j = j0 + lane; j < j1; j += 64) {
        const int k = __builtin_nontemporal_load(col + j);
        const AT a = (AT)__builtin_nontemporal_load(val + j);
        const XT *x = X + (size_t)k * ldx;
#pragma unroll
        for (int v = 0; v < NV; v++) acc[v] += a * (AT)x[v];
    }
#pragma unroll
    for (int v = 0; v < NV; v++) acc[v] = wave_sum(acc[v]);
    if (lane < NV) {
        AT r = 0;
#pragma unroll
        for (int v = 0; v < NV; v++) r = (lane == v) ? acc[v] : r;
        Y[(size_t)row * ldy + lane] = (XT)(r + shift * (AT)X[(size_t)row * ldx + lane]);
    }
}

// Vector-gather kernel: L = next_pow2(ceil(NV / VW)) lanes share one element;
// lane s of the group loads the 16-byte vector X[col, VW*s : VW*s + VW], so a
// wave instruction fetches whole X rows for 64/L elements.  Requires X rows
// 16-byte aligned and readable up to VW*L columns (caller pads).
template <int VW> struct VecT;
template <> struct VecT<4> { typedef float4 type; };
template <> struct VecT<2> { typedef double2 type; };

template <typename VT, typename XT, typename AT, int NV, int VW>
__global__ __launch_bounds__(256) void k_spmm_vec(int nd, const int64_t *__restrict__ rp,
                                                  const int *__restrict__ col, const VT *__restrict__ val,
                                                  const XT *__restrict__ X, int ldx, XT *__restrict__ Y, int ldy,
                                                  AT shift) {
    constexpr int NL = (NV + VW - 1) / VW;                       // lanes with data
    constexpr int L = NL <= 1 ? 1 : NL <= 2 ? 2 : NL <= 4 ? 4 : NL <= 8 ? 8 : NL <= 16 ? 16 : 32;
    constexpr int G = 64 / L;                                    // elements per wave step
    typedef typename VecT<VW>::type V;
    const int lane = threadIdx.x & 63;
    const int g = lane / L, sl = lane % L;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int row = b * 4 + (threadIdx.x >> 6);
    if (row >= nd) return;
    AT acc[VW];
#pragma unroll
    for (int v = 0; v < VW; v++) acc[v] = 0;
    const bool act = sl < NL;
    const int64_t j0 = rp[row], j1 = rp[row + 1];
    for (int64_t j = j0 + g; j < j1; j += G) {
        const int k = __builtin_nontemporal_load(col + j);
        const AT a = (AT)__builtin_nontemporal_load(val + j);
        if (act) {
            const V x = *(const V *)(X + (size_t)k * ldx + sl * VW);
            const XT *xs = (const XT *)&x;
#pragma unroll
            for (int v = 0; v < VW; v++) acc[v] += a * (AT)xs[v];
        }
    }
    // reduce over the G element groups (lane bits above log2(L))
#pragma unroll
    for (int o = 32; o >= L; o >>= 1)
#pragma unroll
        for (int v = 0; v < VW; v++) acc[v] += __shfl_xor(acc[v], o);
    if (g == 0 && act) {
#pragma unroll
        for (int v = 0; v < VW; v++) {
            const int c = sl * VW + v;
            if (c < NV) Y[(size_t)row * ldy + c] = (XT)(acc[v] + shift * (AT)X[(size_t)row * ldx + c]);
        }
    }
}

template <typename VT, typename XT, typename AT>
static void spmm_launch(int nd, const int64_t *rp, const int *col, const VT *val, const XT *X, int ldx, XT *Y,
                        int ldy, int nvec, AT shift, bool padded, hipStream_t s) {
    const unsigned grid = (unsigned)(((nd + 3) / 4 + 7) / 8 * 8);
    constexpr int VW = sizeof(XT) == 4 ? 4 : 2;
    const bool aligned = ((uintptr_t)X % 16 == 0) && ((size_t)ldx * sizeof(XT)) % 16 == 0;
    const bool vec = aligned && (padded || nvec % VW == 0);
    if (vec) {
        switch (nvec) {
#define PCI_NV(n) case n: k_spmm_vec<VT, XT, AT, n, VW><<<grid, 256, 0, s>>>(nd, rp, col, val, X, ldx, Y, ldy, shift); break;
            PCI_NV(1) PCI_NV(2) PCI_NV(3) PCI_NV(4) PCI_NV(5) PCI_NV(6) PCI_NV(7) PCI_NV(8)
            PCI_NV(9) PCI_NV(10) PCI_NV(11) PCI_NV(12) PCI_NV(13) PCI_NV(14) PCI_NV(15) PCI_NV(16)
            PCI_NV(17) PCI_NV(18) PCI_NV(19) PCI_NV(20) PCI_NV(21) PCI_NV(22) PCI_NV(23) PCI_NV(24)
            PCI_NV(28) PCI_NV(32)
#undef PCI_NV
            default: goto scalar;
        }
        PCI_HIP(hipGetLastError());
        return;
    }
scalar:
    switch (nvec) {
#define PCI_NV(n) case n: k_spmm<VT, XT, AT, n><<<grid, 256, 0, s>>>(nd, rp, col, val, X, ldx, Y, ldy, shift); break;
        PCI_NV(1) PCI_NV(2) PCI_NV(3) PCI_NV(4) PCI_NV(5) PCI_NV(6) PCI_NV(7) PCI_NV(8)
        PCI_NV(9) PCI_NV(10) PCI_NV(11) PCI_NV(12) PCI_NV(13) PCI_NV(14) PCI_NV(15) PCI_NV(16)
        PCI_NV(17) PCI_NV(18) PCI_NV(19) PCI_NV(20) PCI_NV(21) PCI_NV(22) PCI_NV(23) PCI_NV(24)
        PCI_NV(25) PCI_NV(26) PCI_NV(27) PCI_NV(28) PCI_NV(29) PCI_NV(30) PCI_NV(31) PCI_NV(32)
        PCI_NV(40) PCI_NV(48) PCI_NV(64)
#undef PCI_NV
        default: throw Error(PCI_E_UNSUPPORTED, "SpMM block width not supported");
    }
    PCI_HIP(hipGetLastError());
}

void spmm_f64(const Ctx &c, const double *X, int ldx, double *Y, int ldy, int nvec, hipStream_t s, bool padded) {
    const Matrix &M = c.H;
    spmm_launch<double, double, double>(M.nd, M.rp.p, M.col.p, M.val.p, X, ldx, Y, ldy, nvec, -M.hmin, padded, s);
}

void spmm_f32(const Ctx &c, const float *X, int ldx, float *Y, int ldy, int nvec, hipStream_t s, bool padded) {
    const Matrix &M = c.H;
    // val32 carries the -hmin shift on its diagonal unless some diagonal
    // element is absent, in which case the shift is applied per row here
    spmm_launch<float, float, float>(M.nd, M.rp.p, M.col.p, M.val32.p, X, ldx, Y, ldy, nvec, M.shift32, padded, s);
}

double spmm_bytes(const Ctx &c, int nvec, int prec) {
    const Matrix &M = c.H;
    const double sv = prec == 32 ? 4.0 : 8.0;
    return (double)M.nnz_full * (sv + 4.0) + 8.0 * (M.nd + 1) + 2.0 * M.nd * nvec * sv;
}

}  // namespace pci

using namespace pci;

extern "C" int pci_gpu_spmm(pci_ctx *ctx, int32_t nvec, int32_t precision, const double *x, double *y,
                            int32_t reps, double *ms, double *bytes) {
    try {
        PCI_REQUIRE(ctx && x && y, PCI_E_ARG, "null argument");
        Ctx &c = ctx->c;
        PCI_HIP(hipSetDevice(c.device));
        PCI_REQUIRE(c.H.valid, PCI_E_ARG, "no Hamiltonian on the device");
        PCI_REQUIRE(nvec >= 1 && nvec <= 64, PCI_E_ARG, "nvec out of range");
        PCI_REQUIRE(precision == 32 || precision == 64, PCI_E_ARG, "precision must be 32 or 64");
        if (reps < 1) reps = 1;
        hipStream_t s = c.stream;
        const int nd = c.H.nd;
        const size_t n = (size_t)nd * nvec;
        hipEvent_t e0, e1;
        PCI_HIP(hipEventCreate(&e0));
        PCI_HIP(hipEventCreate(&e1));
        float t = 0.f;
        // device blocks padded to a multiple of 16 bytes per row
        if (precision == 64) {
            const int ld = (nvec + 1) / 2 * 2;
            std::vector<double> hx((size_t)nd * ld, 0.0);
            for (size_t i = 0; i < (size_t)nd; i++)
                for (int v = 0; v < nvec; v++) hx[i * ld + v] = x[i * nvec + v];
            DBuf<double> dx, dy;
            upload(dx, hx.data(), hx.size(), s);
            dy.alloc(hx.size());
            PCI_HIP(hipEventRecord(e0, s));
            for (int r = 0; r < reps; r++) spmm_f64(c, dx.p, ld, dy.p, ld, nvec, s, true);
            PCI_HIP(hipEventRecord(e1, s));
            PCI_HIP(hipMemcpyAsync(hx.data(), dy.p, hx.size() * sizeof(double), hipMemcpyDeviceToHost, s));
            PCI_HIP(hipStreamSynchronize(s));
            for (size_t i = 0; i < (size_t)nd; i++)
                for (int v = 0; v < nvec; v++) y[i * nvec + v] = hx[i * ld + v];
        } else {
            ensure_fp32(c);
            const int ld = (nvec + 3) / 4 * 4;
            std::vector<float> hx((size_t)nd * ld, 0.f);
            for (size_t i = 0; i < (size_t)nd; i++)
                for (int v = 0; v < nvec; v++) hx[i * ld + v] = (float)x[i * nvec + v];
            DBuf<float> dx, dy;
            upload(dx, hx.data(), hx.size(), s);
            dy.alloc(hx.size());
            PCI_HIP(hipEventRecord(e0, s));
            for (int r = 0; r < reps; r++) spmm_f32(c, dx.p, ld, dy.p, ld, nvec, s, true);
            PCI_HIP(hipEventRecord(e1, s));
            PCI_HIP(hipMemcpyAsync(hx.data(), dy.p, hx.size() * sizeof(float), hipMemcpyDeviceToHost, s));
            PCI_HIP(hipStreamSynchronize(s));
            for (size_t i = 0; i < (size_t)nd; i++)
                for (int v = 0; v < nvec; v++) y[i * nvec + v] = hx[i * ld + v];
        }
        PCI_HIP(hipEventElapsedTime(&t, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (ms) *ms = (double)t / reps;
        if (bytes) *bytes = spmm_bytes(c, nvec, precision);
        return PCI_OK;
    } catch (const pci::Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return PCI_E_HIP;
    }
}
