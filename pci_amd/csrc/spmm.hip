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
// at the target sizes).  Three kernel shapes:
//   k_spmm       one lane per element, nvec scalar X loads (any layout)
//   k_spmm_vec   a group of lanes per element, one 16-byte X load per lane
//   k_spmm_lane  one lane per element, the whole X row in 16-byte loads,
//                U elements per lane in flight (U x 64 per wave step)
//   k_spmm_panel the lane kernel over K column panels (cache blocking): the
//                columns are cut into K ranges so that the X rows one panel
//                gathers fit in an XCD's L2; blocks run panel by panel and
//                write per-panel partial sums, k_spmm_panel_sum adds them in
//                panel order (deterministic) and applies the shift
// Each lane keeps nvec partial sums in registers; a butterfly reduction
// over the wave ends the row.  Rows are assigned so consecutive workgroups
// stay on one XCD (blockIdx round-robins over the 8 XCDs).  The fastest
// shape for a matrix and block width is picked once by timing the
// candidates on the first call (autotuner).
#include <cstdlib>
#include <type_traits>

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



// Dot products of one row segment [j0, j1) with the X rows it references,
// for the lane-per-element kernels.  Software-pipelined: the column/value
// loads of the next step are issued right after the X loads of the current
// step (loads return in issue order, so the FMAs wait for the X gathers
// only), and the next step's matrix stream is in flight meanwhile.  Element
// indices are clamped to the segment, so every load is unconditional and
// in bounds; elements past j1 get the value 0.
template <typename VT, typename XT, typename AT, int NVV, int VW, int U>
__device__ __forceinline__ void row_segment(int64_t j0, int64_t j1, int lane, const int *__restrict__ col,
                                            const VT *__restrict__ val, const XT *__restrict__ X, int ldx,
                                            AT (&acc)[NVV * VW]) {
    typedef typename VecT<VW>::type V;
    if (j1 <= j0) return;  // uniform over the wave
    const int64_t jl = j1 - 1;
    int kc[U];
    AT ac[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int64_t jj = j0 + lane + 64 * u;
        const int64_t jc = jj < jl ? jj : jl;
        kc[u] = __builtin_nontemporal_load(col + jc);
        const AT v = (AT)__builtin_nontemporal_load(val + jc);
        ac[u] = jj < j1 ? v : (AT)0;
    }
    for (int64_t base = j0; base < j1; base += 64 * U) {  // same trip count on every lane
        const int64_t j = base + lane;
        V x[U][NVV];
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int q = 0; q < NVV; q++) x[u][q] = *(const V *)(X + (size_t)kc[u] * ldx + q * VW);
        int kn[U];
        AT an[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t jj = j + 64 * (U + u);
            const int64_t jc = jj < jl ? jj : jl;
            kn[u] = __builtin_nontemporal_load(col + jc);
            const AT v = (AT)__builtin_nontemporal_load(val + jc);
            an[u] = jj < j1 ? v : (AT)0;
        }
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int q = 0; q < NVV; q++) {
                const XT *xs = (const XT *)&x[u][q];
#pragma unroll
                for (int v = 0; v < VW; v++) acc[q * VW + v] += ac[u] * (AT)xs[v];
            }
#pragma unroll
        for (int u = 0; u < U; u++) {
            kc[u] = kn[u];
            ac[u] = an[u];
        }
    }
}

// Lane-per-element kernel with vector X loads: lane l of the wave handles
// elements j0 + l + 64 u (u < U) of each step, so a step keeps 64 U matrix
// elements and their X rows in flight.  Out-of-range elements use column
// `row` with value 0 (branch-free).  Requires X rows 16-byte aligned and
// readable up to NVV * VW columns.
// WPB consecutive rows per workgroup (one per wavefront): with 16 rows on
// one CU the waves walk similar column ranges at the same time, so X rows
// fetched by one wave are often still in that CU's L1 for the others.
template <typename VT, typename XT, typename AT, int NV, int VW, int U, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_spmm_lane(int nd, const int64_t *__restrict__ rp,
                                                        const int *__restrict__ col, const VT *__restrict__ val,
                                                        const XT *__restrict__ X, int ldx, XT *__restrict__ Y,
                                                        int ldy, AT shift) {
    constexpr int NVV = (NV + VW - 1) / VW;
    const int lane = threadIdx.x & 63;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int row = b * WPB + (threadIdx.x >> 6);
    if (row >= nd) return;
    AT acc[NVV * VW];
#pragma unroll
    for (int v = 0; v < NVV * VW; v++) acc[v] = 0;
    const int64_t j0 = rp[row], j1 = rp[row + 1];
    row_segment<VT, XT, AT, NVV, VW, U>(j0, j1, lane, col, val, X, ldx, acc);
#pragma unroll
    for (int v = 0; v < NV; v++) acc[v] = wave_sum(acc[v]);
    if (lane < NV) {
        AT r = 0;
#pragma unroll
        for (int v = 0; v < NV; v++) r = (lane == v) ? acc[v] : r;
        Y[(size_t)row * ldy + lane] = (XT)(r + shift * (AT)X[(size_t)row * ldx + lane]);
    }
}


// Cache-blocked variant: block b works on panel b / row_blocks (blocks are
// dispatched in order, so the chip works through the panels one after the
// other and each XCD's L2 holds one panel of X).  pp: (K+1) x nd offsets.
template <typename VT, typename XT, typename AT, int NV, int VW, int U>
__global__ __launch_bounds__(256) void k_spmm_panel(int nd, const int64_t *__restrict__ pp,
                                                    const int *__restrict__ col, const VT *__restrict__ val,
                                                    const XT *__restrict__ X, int ldx, AT *__restrict__ part) {
    constexpr int NVV = (NV + VW - 1) / VW;
    const int lane = threadIdx.x & 63;
    const int row_blocks = (nd + 3) / 4;
    const int p = blockIdx.x / row_blocks;
    const int row = (blockIdx.x - p * row_blocks) * 4 + (threadIdx.x >> 6);
    if (row >= nd) return;
    AT acc[NVV * VW];
#pragma unroll
    for (int v = 0; v < NVV * VW; v++) acc[v] = 0;
    const int64_t j0 = pp[(size_t)p * nd + row], j1 = pp[(size_t)(p + 1) * nd + row];
    row_segment<VT, XT, AT, NVV, VW, U>(j0, j1, lane, col, val, X, ldx, acc);
#pragma unroll
    for (int v = 0; v < NV; v++) acc[v] = wave_sum(acc[v]);
    if (lane < NV) {
        AT r = 0;
#pragma unroll
        for (int v = 0; v < NV; v++) r = (lane == v) ? acc[v] : r;
        part[((size_t)p * nd + row) * NV + lane] = r;
    }
}

template <typename XT, typename AT>
__global__ __launch_bounds__(256) void k_spmm_panel_sum(int nd, int nv, int K, const AT *__restrict__ part,
                                                        const XT *__restrict__ X, int ldx, XT *__restrict__ Y,
                                                        int ldy, AT shift) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nd * nv) return;
    const int64_t row = i / nv;
    const int v = (int)(i - row * nv);
    AT s = 0;
    for (int p = 0; p < K; p++) s += part[((size_t)p * nd + row) * nv + v];
    Y[(size_t)row * ldy + v] = (XT)(s + shift * (AT)X[(size_t)row * ldx + v]);
}

// pp[p * nd + row] = first element of row with column >= p * ceil(nd / K)
__global__ void k_panel_ptr(int nd, int K, const int64_t *__restrict__ rp, const int *__restrict__ col,
                            int64_t *__restrict__ pp) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= nd) return;
    const int width = (nd + K - 1) / K;
    const int64_t a = rp[row], b = rp[row + 1];
    pp[row] = a;
    pp[(size_t)K * nd + row] = b;
    for (int p = 1; p < K; p++) {
        const int c = p * width;
        int64_t lo = a, hi = b;  // first index in [a, b) with col >= c
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (col[mid] < c) lo = mid + 1;
            else hi = mid;
        }
        pp[(size_t)p * nd + row] = lo;
    }
}


// LDS-staged variant: a workgroup of 16 wavefronts owns R = 16 * RPW
// consecutive rows (RPW per wavefront) and walks the columns in panels of
// P X rows.  Per panel the workgroup copies X[panel] into LDS with coalesced
// loads, then every wavefront consumes, for each of its rows, the elements
// whose column falls in the panel (a contiguous run: columns are ascending),
// 64 at a time, reading X from LDS.  X is fetched from L2 once per
// workgroup and panel instead of once per element, which removes the
// per-element L1 misses that bound the gather kernels.  Each row's current
// chunk of the matrix stream stays in registers across panels, so every
// element is read from memory once.
constexpr int XLDS_BYTES = 65536;

template <typename VT, typename XT, typename AT, int NV, int VW, int RPW, int U>
__global__ __launch_bounds__(1024) void k_spmm_xlds(int nd, const int64_t *__restrict__ rp,
                                                    const int *__restrict__ col, const VT *__restrict__ val,
                                                    const XT *__restrict__ X, int ldx, XT *__restrict__ Y, int ldy,
                                                    AT shift) {
    constexpr int NVV = (NV + VW - 1) / VW;
    constexpr int NVP = NVV * VW;                          // X row width in LDS
    constexpr int P = XLDS_BYTES / (NVP * (int)sizeof(XT));  // X rows per panel
    constexpr int CH = 64 * U;                             // elements per chunk
    typedef typename VecT<VW>::type V;
    __shared__ __align__(16) char smem[XLDS_BYTES];
    V *sXv = (V *)smem;
    const XT *sX = (const XT *)smem;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int row0 = blockIdx.x * (16 * RPW) + wv * RPW;
    // per row: the chunk [jc, jc + CH) held in registers (element jc + lane
    // + 64 u in lane, slot u), the consumption point j inside it, the row end
    int64_t j[RPW], jc[RPW], je[RPW];
    int c[RPW][U];
    AT a[RPW][U];
    auto load_chunk = [&](int k) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t jj = jc[k] + lane + 64 * u;
            const bool ok = jj < je[k];
            c[k][u] = ok ? col[jj] : INT32_MAX;
            a[k][u] = ok ? (AT)val[jj] : (AT)0;
        }
    };
#pragma unroll
    for (int k = 0; k < RPW; k++) {
        const int r = row0 + k;
        jc[k] = j[k] = r < nd ? rp[r] : 0;
        je[k] = r < nd ? rp[r + 1] : 0;
        load_chunk(k);
    }
    AT acc[RPW][NVP];
#pragma unroll
    for (int k = 0; k < RPW; k++)
#pragma unroll
        for (int v = 0; v < NVP; v++) acc[k][v] = 0;
    for (int p0 = 0; p0 < nd; p0 += P) {
        const int np = min(P, nd - p0);
        __syncthreads();  // the previous panel is consumed
        for (int e = tid; e < np * NVV; e += 1024) {
            const int r = e / NVV, q = e - r * NVV;
            sXv[e] = *(const V *)(X + (size_t)(p0 + r) * ldx + q * VW);
        }
        __syncthreads();
        const int pend = p0 + np;
#pragma unroll
        for (int k = 0; k < RPW; k++) {
            while (true) {  // wave-uniform
                int n = 0;
#pragma unroll
                for (int u = 0; u < U; u++) {
                    // not yet consumed and in this panel: a contiguous run from j
                    const bool in = (jc[k] + lane + 64 * u >= j[k]) && c[k][u] < pend;
                    if (in) {
                        const XT *xr = sX + (size_t)(c[k][u] - p0) * NVP;
#pragma unroll
                        for (int q = 0; q < NVV; q++) {
                            const V x = *(const V *)(xr + q * VW);
                            const XT *xs = (const XT *)&x;
#pragma unroll
                            for (int v = 0; v < VW; v++) acc[k][q * VW + v] += a[k][u] * (AT)xs[v];
                        }
                    }
                    n += __popcll(__ballot(in));
                }
                j[k] += n;
                if (j[k] < jc[k] + CH || j[k] >= je[k]) break;  // the rest belongs to later panels
                jc[k] = j[k];
                load_chunk(k);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < RPW; k++) {
        const int r = row0 + k;
#pragma unroll
        for (int v = 0; v < NV; v++) acc[k][v] = wave_sum(acc[k][v]);
        if (r < nd && lane < NV) {
            AT out = 0;
#pragma unroll
            for (int v = 0; v < NV; v++) out = (lane == v) ? acc[k][v] : out;
            Y[(size_t)r * ldy + lane] = (XT)(out + shift * (AT)X[(size_t)r * ldx + lane]);
        }
    }
}


// Pipelined LDS-staged variant: 32 KB panels, the next panel's X loaded into
// registers while the current one is consumed from LDS (stored after the
// workgroup barrier), and a row's next matrix chunk fetched before the
// current chunk is processed whenever the whole current chunk falls in the
// panel.  Same element order and sums as k_spmm_xlds.
constexpr int XLDS2_BYTES = 32768;

template <typename VT, typename XT, typename AT, int NV, int VW, int RPW>
__global__ __launch_bounds__(1024) void k_spmm_xlds2(int nd, const int64_t *__restrict__ rp,
                                                     const int *__restrict__ col, const VT *__restrict__ val,
                                                     const XT *__restrict__ X, int ldx, XT *__restrict__ Y, int ldy,
                                                     AT shift) {
    constexpr int NVV = (NV + VW - 1) / VW;
    constexpr int NVP = NVV * VW;
    constexpr int P = XLDS2_BYTES / (NVP * (int)sizeof(XT));
    constexpr int PV = (P * NVV + 1023) / 1024;  // X vectors per thread and panel
    typedef typename VecT<VW>::type V;
    __shared__ __align__(16) char smem[XLDS2_BYTES];
    V *sXv = (V *)smem;
    const XT *sX = (const XT *)smem;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int row0 = blockIdx.x * (16 * RPW) + wv * RPW;
    int64_t j[RPW], jc[RPW], je[RPW];
    int c[RPW];
    AT a[RPW];
#pragma unroll
    for (int k = 0; k < RPW; k++) {
        const int r = row0 + k;
        jc[k] = j[k] = r < nd ? rp[r] : 0;
        je[k] = r < nd ? rp[r + 1] : 0;
        const int64_t jj = jc[k] + lane;
        const bool ok = jj < je[k];
        c[k] = ok ? col[jj] : INT32_MAX;
        a[k] = ok ? (AT)val[jj] : (AT)0;
    }
    AT acc[RPW][NVP];
#pragma unroll
    for (int k = 0; k < RPW; k++)
#pragma unroll
        for (int v = 0; v < NVP; v++) acc[k][v] = 0;
    V xr[PV];
    auto fetch_panel = [&](int p0) {
        const int np = min(P, nd - p0);
#pragma unroll
        for (int t = 0; t < PV; t++) {
            const int e = tid + t * 1024;
            if (e < np * NVV) {
                const int r = e / NVV, q = e - r * NVV;
                xr[t] = *(const V *)(X + (size_t)(p0 + r) * ldx + q * VW);
            }
        }
    };
    auto store_panel = [&](int p0) {
        const int np = min(P, nd - p0);
#pragma unroll
        for (int t = 0; t < PV; t++) {
            const int e = tid + t * 1024;
            if (e < np * NVV) sXv[e] = xr[t];
        }
    };
    fetch_panel(0);
    store_panel(0);
    __syncthreads();
    for (int p0 = 0; p0 < nd; p0 += P) {
        const int pend = min(p0 + P, nd);
        if (pend < nd) fetch_panel(pend);  // next panel, in flight during this one
#pragma unroll
        for (int k = 0; k < RPW; k++) {
            while (true) {  // wave-uniform
                // the whole chunk is in this panel: fetch the next one first
                const bool whole = __shfl(c[k], 63) < pend && jc[k] + 64 < je[k];
                int cn = INT32_MAX;
                AT an = 0;
                if (whole) {
                    const int64_t jj = jc[k] + 64 + lane;
                    const bool ok = jj < je[k];
                    cn = ok ? col[jj] : INT32_MAX;
                    an = ok ? (AT)val[jj] : (AT)0;
                }
                const bool in = (jc[k] + lane >= j[k]) && c[k] < pend;
                if (in) {
                    const XT *xrow = sX + (size_t)(c[k] - p0) * NVP;
#pragma unroll
                    for (int q = 0; q < NVV; q++) {
                        const V x = *(const V *)(xrow + q * VW);
                        const XT *xs = (const XT *)&x;
#pragma unroll
                        for (int v = 0; v < VW; v++) acc[k][q * VW + v] += a[k] * (AT)xs[v];
                    }
                }
                j[k] += __popcll(__ballot(in));
                if (whole) {  // j reached jc + 64: move to the prefetched chunk
                    jc[k] += 64;
                    c[k] = cn;
                    a[k] = an;
                    continue;
                }
                if (j[k] < jc[k] + 64 || j[k] >= je[k]) break;
                jc[k] = j[k];  // chunk used up exactly at the end of the row part in this panel
                const int64_t jj = jc[k] + lane;
                const bool ok = jj < je[k];
                c[k] = ok ? col[jj] : INT32_MAX;
                a[k] = ok ? (AT)val[jj] : (AT)0;
            }
        }
        __syncthreads();  // LDS panel consumed by every wave
        if (pend < nd) {
            store_panel(pend);
            __syncthreads();
        }
    }
#pragma unroll
    for (int k = 0; k < RPW; k++) {
        const int r = row0 + k;
#pragma unroll
        for (int v = 0; v < NV; v++) acc[k][v] = wave_sum(acc[k][v]);
        if (r < nd && lane < NV) {
            AT out = 0;
#pragma unroll
            for (int v = 0; v < NV; v++) out = (lane == v) ? acc[k][v] : out;
            Y[(size_t)r * ldy + lane] = (XT)(out + shift * (AT)X[(size_t)r * ldx + lane]);
        }
    }
}

// LDS-staged variant with several rows per wavefront at once: the wave is
// split into GR groups of L = 64 / GR lanes, group g working on its own row
// (L elements per step), and the wave walks RQ such row sets, so a
// workgroup of 16 waves owns 16 * GR * RQ rows that share every X panel
// staged in LDS.  Needs NV <= L (the group's lanes hold the row's sums).
template <typename VT, typename XT, typename AT, int NV, int VW, int GR, int RQ>
__global__ __launch_bounds__(1024) void k_spmm_xldsg(int nd, const int64_t *__restrict__ rp,
                                                     const int *__restrict__ col, const VT *__restrict__ val,
                                                     const XT *__restrict__ X, int ldx, XT *__restrict__ Y, int ldy,
                                                     AT shift) {
    constexpr int L = 64 / GR;
    constexpr int NVV = (NV + VW - 1) / VW;
    constexpr int NVP = NVV * VW;
    constexpr int P = XLDS_BYTES / (NVP * (int)sizeof(XT));
    typedef typename VecT<VW>::type V;
    __shared__ __align__(16) char smem[XLDS_BYTES];
    V *sXv = (V *)smem;
    const XT *sX = (const XT *)smem;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int g = lane / L, gl = lane % L;
    const uint64_t gmask = (L == 64 ? ~0ull : ((1ull << L) - 1)) << (g * L);
    // rows of this lane's group: row0 + q * GR + g, q < RQ
    const int row0 = (blockIdx.x * 16 + wv) * (GR * RQ);
    int64_t j[RQ], je[RQ];
#pragma unroll
    for (int q = 0; q < RQ; q++) {
        const int r = row0 + q * GR + g;
        j[q] = r < nd ? rp[r] : 0;
        je[q] = r < nd ? rp[r + 1] : 0;
    }
    AT acc[RQ][NVP];
#pragma unroll
    for (int q = 0; q < RQ; q++)
#pragma unroll
        for (int v = 0; v < NVP; v++) acc[q][v] = 0;
    for (int p0 = 0; p0 < nd; p0 += P) {
        const int np = min(P, nd - p0);
        __syncthreads();
        for (int e = tid; e < np * NVV; e += 1024) {
            const int r = e / NVV, qq = e - r * NVV;
            sXv[e] = *(const V *)(X + (size_t)(p0 + r) * ldx + qq * VW);
        }
        __syncthreads();
        const int pend = p0 + np;
#pragma unroll
        for (int q = 0; q < RQ; q++) {
            bool active = j[q] < je[q];
            while (__ballot(active)) {  // wave-uniform
                bool in = false;
                if (active) {
                    const int64_t jj = j[q] + gl;
                    const bool ok = jj < je[q];
                    const int c = ok ? col[jj] : INT32_MAX;
                    const AT a = ok ? (AT)val[jj] : (AT)0;
                    in = c < pend;
                    if (in) {
                        const XT *xr = sX + (size_t)(c - p0) * NVP;
#pragma unroll
                        for (int t = 0; t < NVV; t++) {
                            const V x = *(const V *)(xr + t * VW);
                            const XT *xs = (const XT *)&x;
#pragma unroll
                            for (int v = 0; v < VW; v++) acc[q][t * VW + v] += a * (AT)xs[v];
                        }
                    }
                }
                const int n = __popcll(__ballot(in) & gmask);
                if (active) {
                    j[q] += n;
                    if (n < L || j[q] >= je[q]) active = false;
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < RQ; q++) {
#pragma unroll
        for (int v = 0; v < NV; v++)
#pragma unroll
            for (int o = L / 2; o > 0; o >>= 1) acc[q][v] += __shfl_xor(acc[q][v], o);
        const int r = row0 + q * GR + g;
        if (r < nd && gl < NV) {
            AT out = 0;
#pragma unroll
            for (int v = 0; v < NV; v++) out = (gl == v) ? acc[q][v] : out;
            Y[(size_t)r * ldy + gl] = (XT)(out + shift * (AT)X[(size_t)r * ldx + gl]);
        }
    }
}

// ---------------------------------------------------------------- sliced copy
// LDS-panel kernel over the sliced copy (SellCopy).  A workgroup of W
// wavefronts owns W chunks of 16 rows (a row block) and a range of pg column
// panels (a tile: row block x panel group).  Per panel it copies the panel's
// X rows into LDS, then each wavefront streams its chunk's segment for the
// panel: every load of the matrix is one contiguous 256- or 512-byte run per
// wavefront (lane-interleaved layout), each lane accumulates the partial
// sums of its row for all nv columns in registers, and nothing is reduced
// until the tile ends (4 lanes per row, two butterfly steps).  Because a row
// block spans 16 * W rows, one X panel fill serves that many rows' elements,
// against 32 rows in k_spmm_xlds.  With more than
// one panel group the tiles write per-group partial sums, added in group
// order by k_spmm_panel_sum (deterministic).  The matrix stream is read with
// non-temporal loads so that it does not evict X from L2.
constexpr int SELL_R = 16;  // rows per chunk (one wavefront)
constexpr int SELL_L = 4;   // lanes per row

constexpr int SELL_LDS = 65536;  // bytes of X panel per workgroup

// X row stride in the LDS panel, in values, for a block of nv columns: nv
// padded to the 16-byte read width, plus one 16-byte vector where the row
// would span an even number (above 2) of vectors.  The lanes of a wavefront
// read rows at random; a stride of s vectors puts every row's start on one of
// 16 / gcd(s, 16) groups of 4 of the 64 LDS banks (2 for s = 8), and the
// fewer the groups, the more those reads conflict.  An odd stride reaches all
// 16; s = 2 keeps 8 and its wider panels.
template <typename XT>
constexpr int sell_ldx(int nv) {
    constexpr int VW = 16 / (int)sizeof(XT);
    const int nvv = (nv + VW - 1) / VW;
    return (nvv > 2 && nvv % 2 == 0 ? nvv + 1 : nvv) * VW;
}

// panel width for a block of nv columns: as many X rows as fit the LDS
// panel, at most 65536 (16-bit panel-local columns)
template <typename XT>
inline int sell_panel(int nv) {
    return std::min(65536, SELL_LDS / (sell_ldx<XT>(nv) * (int)sizeof(XT)));
}

// rs[row * (npan + 1) + p] = offset, from the row start, of the row's first
// element with column >= p * P (columns ascend within a row)
__global__ void k_sell_runs(int nd, int npan, int P, const int64_t *__restrict__ rp, const int *__restrict__ col,
                            int *__restrict__ rs) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nd * (npan + 1)) return;
    const int row = (int)(t / (npan + 1)), p = (int)(t - (int64_t)row * (npan + 1));
    const int64_t a = rp[row], b = rp[row + 1];
    int64_t lo = a, hi = b;
    if (p == npan) {
        lo = b;
    } else {
        const int c = p * P;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (col[mid] < c) lo = mid + 1;
            else hi = mid;
        }
    }
    rs[t] = (int)(lo - a);
}

// pair steps of segment (c, p): the longest run among the chunk's rows,
// in steps of 4 elements (one per lane of the row), rounded up to pairs
__global__ void k_sell_depth(int nd, int npan, int64_t nseg, const int *__restrict__ rs,
                             int64_t *__restrict__ depth) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nseg) return;
    const int c = (int)(t / npan), p = (int)(t - (int64_t)c * npan);
    int m = 0;
    for (int r = c * SELL_R; r < min(nd, (c + 1) * SELL_R); r++) {
        const int *q = rs + (size_t)r * (npan + 1) + p;
        m = max(m, q[1] - q[0]);
    }
    const int steps = (m + SELL_L - 1) / SELL_L;
    depth[t] = (steps + 1) >> 1;
}

// one wavefront per row: element j of row r, k-th of its run in panel p,
// goes to lane 4 (r mod 16) + k mod 4, step k / 4 of segment (r / 16, p)
template <typename VT>
__global__ __launch_bounds__(256) void k_sell_fill(int nd, int npan, int P, const int64_t *__restrict__ rp,
                                                   const int *__restrict__ col, const VT *__restrict__ val,
                                                   const int *__restrict__ rs, const int64_t *__restrict__ off,
                                                   uint16_t *__restrict__ scol, VT *__restrict__ sval) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= nd) return;
    const int64_t a = rp[row], b = rp[row + 1];
    const int *rr = rs + (size_t)row * (npan + 1);
    const int64_t segrow = (int64_t)(row / SELL_R) * npan;
    const int rl = (row % SELL_R) * SELL_L;
    for (int64_t j = a + lane; j < b; j += 64) {
        const int c = col[j];
        const int p = c / P;
        const int k = (int)(j - a) - rr[p];
        const int step = k / SELL_L;
        const int64_t idx = (off[segrow + p] + (step >> 1)) * 128 + (int64_t)(rl + k % SELL_L) * 2 + (step & 1);
        scol[idx] = (uint16_t)(c - p * P);
        sval[idx] = val[j];
    }
}

template <typename VT>
static void sell_build_impl(Matrix &M, SellCopy &S, const VT *val, int P, hipStream_t s, DBuf<char> &work);

// Builds the sliced copy, or marks it ineligible (the callers then use the
// CSR kernels) when it does not fit the device memory left: the copy is
// checked against hipMemGetInfo with a 2 GB margin once its size is known,
// and an allocation failure releases whatever part was built.
// PCI_SELL_OFF=1 marks every copy ineligible (tests of the fallback).
template <typename VT>
static void sell_build(Matrix &M, SellCopy &S, const VT *val, int P, hipStream_t s, DBuf<char> &work) {
    const char *off = getenv("PCI_SELL_OFF");
    if (off && atoi(off) != 0) {
        S.eligible = false;
        return;
    }
    try {
        sell_build_impl<VT>(M, S, val, P, s, work);
    } catch (const Error &e) {
        if (e.code != PCI_E_NOMEM) throw;
        (void)hipGetLastError();
        (void)hipStreamSynchronize(s);
        S.release();
        S.eligible = false;
    }
}

template <typename VT>
static void sell_build_impl(Matrix &M, SellCopy &S, const VT *val, int P, hipStream_t s, DBuf<char> &work) {
    const int nd = M.nd;
    S.P = P;
    S.npan = (nd + P - 1) / P;
    S.nchunk = (nd + SELL_R - 1) / SELL_R;
    const int64_t nrs = (int64_t)nd * (S.npan + 1), nseg = S.nchunk * S.npan;
    // beyond this the run table alone outgrows its use (X panels would be a
    // small fraction of the row's columns): leave the copy to larger panels
    if (nrs > ((int64_t)1 << 30) || nd < 64) {
        S.eligible = false;
        return;
    }
    hipEvent_t e0, e1;
    PCI_HIP(hipEventCreate(&e0));
    PCI_HIP(hipEventCreate(&e1));
    PCI_HIP(hipEventRecord(e0, s));
    DBuf<int> rs;
    rs.alloc((size_t)nrs);
    k_sell_runs<<<(unsigned)((nrs + 255) / 256), 256, 0, s>>>(nd, S.npan, P, M.rp.p, M.col.p, rs.p);
    DBuf<int64_t> depth;
    depth.alloc((size_t)nseg);
    k_sell_depth<<<(unsigned)((nseg + 255) / 256), 256, 0, s>>>(nd, S.npan, nseg, rs.p, depth.p);
    S.off.alloc((size_t)nseg + 1);
    scan_exclusive(depth.p, S.off.p, nseg, s, work);
    int64_t pairs = 0;
    PCI_HIP(hipMemcpyAsync(&pairs, S.off.p + nseg, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    PCI_HIP(hipStreamSynchronize(s));
    S.pairs = pairs;
    {
        size_t fr = 0, tot = 0;
        PCI_HIP(hipMemGetInfo(&fr, &tot));
        const size_t need = (size_t)pairs * 64 * sizeof(uint32_t) + (size_t)pairs * 128 * sizeof(VT);
        if (need + ((size_t)2 << 30) > fr)
            throw Error(PCI_E_NOMEM, "sliced copy of " + std::to_string(need) + " bytes does not fit");
    }
    S.col.alloc((size_t)pairs * 64);
    S.val.alloc((size_t)pairs * 128 * sizeof(VT));
    PCI_HIP(hipMemsetAsync(S.col.p, 0, S.col.bytes(), s));
    PCI_HIP(hipMemsetAsync(S.val.p, 0, S.val.bytes(), s));
    k_sell_fill<VT><<<(unsigned)((nd + 3) / 4), 256, 0, s>>>(nd, S.npan, P, M.rp.p, M.col.p, val, rs.p, S.off.p,
                                                             (uint16_t *)S.col.p, (VT *)S.val.p);
    PCI_HIP(hipGetLastError());
    PCI_HIP(hipEventRecord(e1, s));
    PCI_HIP(hipEventSynchronize(e1));
    PCI_HIP(hipEventElapsedTime(&S.build_ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    S.built = true;
}

template <typename VT, int N>
struct VecN;
template <> struct VecN<float, 2> { typedef float type __attribute__((ext_vector_type(2))); };
template <> struct VecN<double, 2> { typedef double type __attribute__((ext_vector_type(2))); };

template <typename VT, typename XT, typename AT, int NV, int VW, int W, int KC, int U, bool SKIP>
__global__ __launch_bounds__(64 * W) void k_spmm_sell(int nd, int P, int npan, int pg, int nrb,
                                                      const int64_t *__restrict__ off,
                                                      const uint32_t *__restrict__ scol,
                                                      const VT *__restrict__ sval, const XT *__restrict__ X,
                                                      int ldx, XT *__restrict__ Y, int ldy, AT *__restrict__ part,
                                                      AT shift) {
    constexpr int NVV = (NV + VW - 1) / VW;
    constexpr int NVP = NVV * VW;            // NV padded to the read width VW
    constexpr int LS = sell_ldx<XT>(NV);     // X row stride in LDS
    static_assert(VW * sizeof(XT) == 16, "16-byte X reads");
    typedef XT V __attribute__((ext_vector_type(VW)));
    typedef XT V2X __attribute__((ext_vector_type(2)));
    typedef typename VecN<VT, 2>::type V2;
    extern __shared__ __align__(16) char sell_lds[];  // P * LS values of X
    XT *sX = (XT *)sell_lds;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int rb = blockIdx.x % nrb, sg = blockIdx.x / nrb;
    const int p_lo = sg * pg, p_hi = min(npan, p_lo + pg);
    const int nchunk = (nd + SELL_R - 1) / SELL_R;
    const int c0 = (rb * W + wv) * KC;  // this wavefront's first chunk (KC consecutive chunks)
    const V2 *sv = (const V2 *)sval;
    AT acc[KC][NVP];
#pragma unroll
    for (int k = 0; k < KC; k++)
#pragma unroll
        for (int v = 0; v < NVP; v++) acc[k][v] = 0;
    for (int p = p_lo; p < p_hi; p++) {
        const int x0 = p * P, nx = min(P, nd - x0);
        __syncthreads();  // the previous panel is consumed
        for (int e = tid; e < nx * NVV; e += 64 * W) {
            const int r = e / NVV, q = e - r * NVV;
            *(V *)(sX + r * LS + q * VW) = *(const V *)(X + (size_t)(x0 + r) * ldx + q * VW);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KC; k++) {
            const int c = c0 + k;
            if (c >= nchunk) break;  // wave-uniform
            const int64_t seg = (int64_t)c * npan + p;
            const int64_t e = off[seg + 1];
            for (int64_t b = off[seg]; b < e; b += U) {
                uint32_t cc[U];
                V2 vv[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    if (b + u < e) {
                        cc[u] = __builtin_nontemporal_load(scol + (b + u) * 64 + lane);
                        vv[u] = __builtin_nontemporal_load(sv + (b + u) * 64 + lane);
                    } else {
                        cc[u] = 0;
                        vv[u] = (V2)(VT)0;
                    }
                }
#pragma unroll
                for (int u = 0; u < U; u++) {
#pragma unroll
                    for (int h = 0; h < 2; h++) {
                        const int cl = h ? (int)(cc[u] >> 16) : (int)(cc[u] & 0xffffu);
                        const AT a = (AT)vv[u][h];
                        // padding slots hold 0: with SKIP their lanes read no X
                        if (!SKIP || a != (AT)0) {
                            const XT *xr = sX + cl * LS;
#pragma unroll
                            for (int q = 0; q < NVV; q++) {
                                // the last vector reads only its real columns when
                                // one narrower read covers them
                                if (NV - q * VW == 1) {
                                    acc[k][q * VW] += a * (AT)xr[q * VW];
                                } else if (NV - q * VW == 2 && VW == 4) {
                                    const V2X x = *(const V2X *)(xr + q * VW);
                                    acc[k][q * VW] += a * (AT)x[0];
                                    acc[k][q * VW + 1] += a * (AT)x[1];
                                } else {  // three columns: one 16-byte read beats 8 + 4 bytes
                                    const V x = *(const V *)(xr + q * VW);
#pragma unroll
                                    for (int v = 0; v < VW; v++)
                                        if (q * VW + v < NV) acc[k][q * VW + v] += a * (AT)x[v];
                                }
                            }
                        }
                    }
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < KC; k++) {
#pragma unroll
        for (int v = 0; v < NV; v++) {
            acc[k][v] += __shfl_xor(acc[k][v], 1);
            acc[k][v] += __shfl_xor(acc[k][v], 2);
        }
        const int row = (c0 + k) * SELL_R + (lane >> 2);
        if ((lane & 3) == 0 && row < nd) {
            if (part) {
#pragma unroll
                for (int v = 0; v < NV; v++) part[((size_t)sg * nd + row) * NV + v] = acc[k][v];
            } else {
#pragma unroll
                for (int v = 0; v < NV; v++)
                    Y[(size_t)row * ldy + v] = (XT)(acc[k][v] + shift * (AT)X[(size_t)row * ldx + v]);
            }
        }
    }
}

// Pipelined form of k_spmm_sell for matrices whose rows hold few elements
// per panel (Nd in the millions: a 2048-column panel then holds ~2 pair
// steps of a chunk).  k_spmm_sell walks a wave's KC chunks one after the
// other, each load batch waiting for the previous chunk's FMAs, and starts
// a panel's loads only after its fill barrier, so a wave has one short
// batch in flight at a time.  Here the first U pair steps of every chunk of
// the wave for the next panel are issued together (KC x U loads in flight)
// before the wave waits at that panel's fill barrier -- with PF, before it
// computes the current panel, holding a second set of registers -- and only
// the steps past U are streamed after them.
template <typename VT, typename XT, typename AT, int NV, int VW, int W, int KC, int U, bool PF>
__global__ __launch_bounds__(64 * W) void k_spmm_sellq(int nd, int P, int npan, int pg, int nrb,
                                                       const int64_t *__restrict__ off,
                                                       const uint32_t *__restrict__ scol,
                                                       const VT *__restrict__ sval, const XT *__restrict__ X,
                                                       int ldx, XT *__restrict__ Y, int ldy, AT *__restrict__ part,
                                                       AT shift) {
    constexpr int NVV = (NV + VW - 1) / VW;
    constexpr int NVP = NVV * VW;
    constexpr int LS = sell_ldx<XT>(NV);
    static_assert(VW * sizeof(XT) == 16, "16-byte X reads");
    typedef XT V __attribute__((ext_vector_type(VW)));
    typedef typename VecN<VT, 2>::type V2;
    extern __shared__ __align__(16) char sellq_lds[];
    XT *sX = (XT *)sellq_lds;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int rb = blockIdx.x % nrb, sg = blockIdx.x / nrb;
    const int p_lo = sg * pg, p_hi = min(npan, p_lo + pg);
    const int nchunk = (nd + SELL_R - 1) / SELL_R;
    const int c0 = (rb * W + wv) * KC;
    const V2 *sv = (const V2 *)sval;
    AT acc[KC][NVP];
#pragma unroll
    for (int k = 0; k < KC; k++)
#pragma unroll
        for (int v = 0; v < NVP; v++) acc[k][v] = 0;
    // queue of the first U pair steps of every chunk for one panel
    auto issue = [&](int p, uint32_t (&qc)[KC][U], V2 (&qv)[KC][U]) {
#pragma unroll
        for (int k = 0; k < KC; k++) {
            int64_t b = 0, e = 0;
            if (c0 + k < nchunk && p < p_hi) {
                const int64_t seg = (int64_t)(c0 + k) * npan + p;
                b = off[seg];
                e = off[seg + 1];
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (b + u < e) {
                    qc[k][u] = __builtin_nontemporal_load(scol + (b + u) * 64 + lane);
                    qv[k][u] = __builtin_nontemporal_load(sv + (b + u) * 64 + lane);
                } else {
                    qc[k][u] = 0;
                    qv[k][u] = (V2)(VT)0;
                }
            }
        }
    };
    // one pair step: padding slots hold 0 and read no X
    auto step = [&](int k, uint32_t cc, V2 vv) {
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int cl = h ? (int)(cc >> 16) : (int)(cc & 0xffffu);
            const AT a = (AT)vv[h];
            if (a != (AT)0) {
                const XT *xr = sX + cl * LS;
#pragma unroll
                for (int q = 0; q < NVV; q++) {
                    const V x = *(const V *)(xr + q * VW);
#pragma unroll
                    for (int v = 0; v < VW; v++)
                        if (q * VW + v < NV) acc[k][q * VW + v] += a * (AT)x[v];
                }
            }
        }
    };
    uint32_t qc[KC][U];
    V2 qv[KC][U];
    issue(p_lo, qc, qv);
    for (int p = p_lo; p < p_hi; p++) {
        const int x0 = p * P, nx = min(P, nd - x0);
        __syncthreads();  // the previous panel is consumed
        for (int e = tid; e < nx * NVV; e += 64 * W) {
            const int r = e / NVV, q = e - r * NVV;
            *(V *)(sX + r * LS + q * VW) = *(const V *)(X + (size_t)(x0 + r) * ldx + q * VW);
        }
        __syncthreads();
        uint32_t nc[KC][U];
        V2 nv[KC][U];
        if (PF) issue(p + 1, nc, nv);
#pragma unroll
        for (int k = 0; k < KC; k++) {
#pragma unroll
            for (int u = 0; u < U; u++) step(k, qc[k][u], qv[k][u]);
            if (c0 + k >= nchunk) continue;  // wave-uniform
            const int64_t seg = (int64_t)(c0 + k) * npan + p;
            const int64_t e = off[seg + 1];
            for (int64_t b = off[seg] + U; b < e; b += U) {  // steps past the queue (long segments)
                uint32_t cc[U];
                V2 vv[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    if (b + u < e) {
                        cc[u] = __builtin_nontemporal_load(scol + (b + u) * 64 + lane);
                        vv[u] = __builtin_nontemporal_load(sv + (b + u) * 64 + lane);
                    } else {
                        cc[u] = 0;
                        vv[u] = (V2)(VT)0;
                    }
                }
#pragma unroll
                for (int u = 0; u < U; u++) step(k, cc[u], vv[u]);
            }
        }
        if (PF) {
#pragma unroll
            for (int k = 0; k < KC; k++)
#pragma unroll
                for (int u = 0; u < U; u++) {
                    qc[k][u] = nc[k][u];
                    qv[k][u] = nv[k][u];
                }
        } else {
            issue(p + 1, qc, qv);  // in flight while the wave waits for the next fill
        }
    }
#pragma unroll
    for (int k = 0; k < KC; k++) {
#pragma unroll
        for (int v = 0; v < NV; v++) {
            acc[k][v] += __shfl_xor(acc[k][v], 1);
            acc[k][v] += __shfl_xor(acc[k][v], 2);
        }
        const int row = (c0 + k) * SELL_R + (lane >> 2);
        if ((lane & 3) == 0 && row < nd) {
            if (part) {
#pragma unroll
                for (int v = 0; v < NV; v++) part[((size_t)sg * nd + row) * NV + v] = acc[k][v];
            } else {
#pragma unroll
                for (int v = 0; v < NV; v++)
                    Y[(size_t)row * ldy + v] = (XT)(acc[k][v] + shift * (AT)X[(size_t)row * ldx + v]);
            }
        }
    }
}

// ------------------------------------------------------- residual blocks
// (RblkCopy in internal.h.)  The SpMM over the blocks: a workgroup takes one
// unit (a list r and a range of its 16-row chunks, strided over the W
// waves), gathers X rows of the list's determinants into LDS, column tile
// by column tile of TC, and each wave streams its chunks' values for the
// tile: one 512-byte coalesced load per wavefront and step of 8 columns,
// values only.  k_spmm_rblk multiplies on the vector ALUs: lane l holds
// row l mod 16 and columns l / 16 and 4 + l / 16 of the step, the 16 lanes of
// a column read the same X row (broadcast LDS reads), the row's partial sums
// stay in registers and a 4-lane butterfly (lanes 16 apart) ends them.
// k_spmm_rbmf puts the same tiles on the matrix cores (below).  Either way
// the row sums go to the row's list slot (list_ptr[r] + position).  k_rb_combine then adds, per
// determinant, its Ne(Ne-1)/2 list slots in a fixed order (deterministic)
// and applies the diagonal shift.
// det_pos[i * nres + q] = position of determinant i in its q-th list
__global__ __launch_bounds__(64) void k_rb_pos(ResView R, int *__restrict__ det_pos) {
    const int lid = blockIdx.x;
    const int64_t a = R.list_ptr[lid], b = R.list_ptr[lid + 1];
    for (int64_t k = a + threadIdx.x; k < b; k += 64) {
        const int det = R.list_det[k];
        for (int q = 0; q < R.nres; q++)
            if (R.det_res[(size_t)det * R.nres + q] == lid) det_pos[(size_t)det * R.nres + q] = (int)(k - a);
    }
}

// combo[mask] = index of the rsize-subset `mask` of the Ne positions of a
// determinant in the lexicographic order the residual lists are built with
// (-1 for masks of another size)
constexpr int RB_MAXNE = 8;

// Index of the list a stored element (i, k) belongs to in each determinant,
// from the lowest rsize positions the two share: returns false when they
// share fewer (an element outside the space's coupling structure).
template <int W>
__device__ inline bool rb_canon(const uint64_t (&A)[W], const uint64_t (&B)[W], int rsize, const int *__restrict__ combo,
                                int &qa, int &qb) {
    uint64_t c[W];
#pragma unroll
    for (int w = 0; w < W; w++) c[w] = A[w] & B[w];
    int ma = 0, mb = 0, got = 0;
#pragma unroll
    for (int w = 0; w < W; w++) {
        while (c[w] && got < rsize) {
            const int bit = __builtin_ctzll(c[w]);
            c[w] &= c[w] - 1;
            const uint64_t below = (1ull << bit) - 1ull;
            int ra = popc64(A[w] & below), rb = popc64(B[w] & below);
#pragma unroll
            for (int v = 0; v < W; v++)
                if (v < w) {
                    ra += popc64(A[v]);
                    rb += popc64(B[v]);
                }
            ma |= 1 << ra;
            mb |= 1 << rb;
            got++;
        }
    }
    if (got < rsize) return false;
    qa = combo[ma];
    qb = combo[mb];
    return true;
}

// One wavefront per CSR row: each element goes to its canonical list's block.
template <typename VT, int W>
__global__ __launch_bounds__(256) void k_rb_fill(int nd, ResView R, const int *__restrict__ combo,
                                                 const int *__restrict__ det_pos, const int64_t *__restrict__ rp,
                                                 const int *__restrict__ col, const VT *__restrict__ val,
                                                 const int64_t *__restrict__ boff, VT *__restrict__ out,
                                                 int *__restrict__ bad) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= nd) return;
    uint64_t A[W];
#pragma unroll
    for (int w = 0; w < W; w++) A[w] = R.occ[(size_t)w * nd + row];
    for (int64_t j = rp[row] + lane; j < rp[row + 1]; j += 64) {
        const int k = col[j];
        uint64_t B[W];
#pragma unroll
        for (int w = 0; w < W; w++) B[w] = R.occ[(size_t)w * nd + k];
        int qa = 0, qb = 0;
        if (!rb_canon<W>(A, B, R.rsize, combo, qa, qb)) {
            atomicOr(bad, 1);
            continue;
        }
        const int lid = R.det_res[(size_t)row * R.nres + qa];
        if (R.det_res[(size_t)k * R.nres + qb] != lid) {
            atomicOr(bad, 2);
            continue;
        }
        const int p = det_pos[(size_t)row * R.nres + qa], q = det_pos[(size_t)k * R.nres + qb];
        const int n = (int)(R.list_ptr[lid + 1] - R.list_ptr[lid]);
        out[boff[lid] + rb_index(p, q, (n + 7) >> 3)] = val[j];
    }
}

// KN chunks' worth of UN column steps from step s of the current X tile:
// each X row pair read from LDS serves every chunk (KN, UN compile-time, so
// no multiply-add is spent on a chunk the wave does not have or on a step
// past the tile).  Per chunk the sums run over (step, h, column) in the same
// order whatever KN and UN are, so the results do not depend on them.
template <typename VT, typename XT, typename AT, int NV, int VW, int KC, int KN, int UN>
__device__ __forceinline__ void rb_steps(AT (&acc)[KC][(NV + VW - 1) / VW * VW],
                                         const typename VecN<VT, 2>::type *const (&pk)[KC], int s, int s0, int sub,
                                         const XT *sX) {
    constexpr int NVV = (NV + VW - 1) / VW;
    constexpr int NVP = NVV * VW;
    typedef XT V __attribute__((ext_vector_type(VW)));
    typedef typename VecN<VT, 2>::type V2;
    V2 vv[KN][UN];
#pragma unroll
    for (int k = 0; k < KN; k++)
#pragma unroll
        for (int u = 0; u < UN; u++) vv[k][u] = __builtin_nontemporal_load(pk[k] + (size_t)(s + u) * 64);
#pragma unroll
    for (int u = 0; u < UN; u++) {
        // columns 8 (s+u) + 4 h + sub of the tile
        const XT *xr = sX + (8 * (s + u - s0) + sub) * NVP;
#pragma unroll
        for (int h = 0; h < 2; h++) {
#pragma unroll
            for (int q = 0; q < NVV; q++) {
                const V x = *(const V *)(xr + 4 * h * NVP + q * VW);
#pragma unroll
                for (int k = 0; k < KN; k++) {
                    const AT a = (AT)vv[k][u][h];
#pragma unroll
                    for (int v = 0; v < VW; v++)
                        if (q * VW + v < NV) acc[k][q * VW + v] += a * (AT)x[v];
                }
            }
        }
    }
}

// The tile's steps [s0, s1) for a wave holding KN chunks: U steps at a time,
// then the rest one by one.
template <typename VT, typename XT, typename AT, int NV, int VW, int KC, int KN, int U>
__device__ __forceinline__ void rb_tile(AT (&acc)[KC][(NV + VW - 1) / VW * VW],
                                        const typename VecN<VT, 2>::type *const (&pk)[KC], int s0, int s1, int sub,
                                        const XT *sX) {
    int s = s0;
    for (; s + U <= s1; s += U) rb_steps<VT, XT, AT, NV, VW, KC, KN, U>(acc, pk, s, s0, sub, sX);
    for (; s < s1; s++) rb_steps<VT, XT, AT, NV, VW, KC, KN, 1>(acc, pk, s, s0, sub, sX);
}

template <typename VT, typename XT, typename AT, int NV, int VW, int W, int KC, int U>
__global__ __launch_bounds__(64 * W) void k_spmm_rblk(const int *__restrict__ unit, const int64_t *__restrict__ list_ptr,
                                                      const int *__restrict__ list_det,
                                                      const int64_t *__restrict__ boff, const VT *__restrict__ val,
                                                      int TC, const XT *__restrict__ X, int ldx,
                                                      AT *__restrict__ slot) {
    constexpr int NVV = (NV + VW - 1) / VW;
    constexpr int NVP = NVV * VW;
    static_assert(VW * sizeof(XT) == 16, "16-byte X reads");
    static_assert(KC >= 1 && KC <= 4, "chunks per wave");
    typedef XT V __attribute__((ext_vector_type(VW)));
    typedef typename VecN<VT, 2>::type V2;
    extern __shared__ __align__(16) char rb_lds[];
    XT *sX = (XT *)rb_lds;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int lid = unit[3 * blockIdx.x], ch0 = unit[3 * blockIdx.x + 1], ch1 = unit[3 * blockIdx.x + 2];
    const int64_t lp = list_ptr[lid];
    const int n = (int)(list_ptr[lid + 1] - lp);
    const int steps = (n + 7) >> 3;
    const V2 *sv = (const V2 *)(val + boff[lid]);
    const int sub = lane >> 4;  // the lane's row is lane mod 16, its columns sub and 4 + sub of each step
    // the wave's chunks ch0 + wv + W k, k < kn (wave-uniform); every chunk
    // walks the same column steps of the list's block
    const int kn = ch1 - ch0 - wv <= 0 ? 0 : min(KC, (ch1 - ch0 - wv + W - 1) / W);
    const V2 *pk[KC];
#pragma unroll
    for (int k = 0; k < KC; k++) pk[k] = sv + (size_t)(ch0 + wv + W * min(k, max(kn - 1, 0))) * steps * 64 + lane;
    AT acc[KC][NVP];
#pragma unroll
    for (int k = 0; k < KC; k++)
#pragma unroll
        for (int v = 0; v < NVP; v++) acc[k][v] = 0;
    for (int j0 = 0; j0 < n; j0 += TC) {
        const int nj = min(TC, n - j0);
        const int njp = (nj + 7) & ~7;  // the tile's rows, padded to whole steps with zeros
        __syncthreads();  // the previous tile is consumed
        for (int e = tid; e < njp * NVV; e += 64 * W) {
            const int r = e / NVV, q = e - r * NVV;
            V x = (V)(XT)0;
            if (r < nj) x = *(const V *)(X + (size_t)list_det[lp + j0 + r] * ldx + q * VW);
            *(V *)(sX + r * NVP + q * VW) = x;
        }
        __syncthreads();
        const int s0 = j0 >> 3, s1 = s0 + (njp >> 3);
        switch (kn) {  // wave-uniform
            case 0: break;
            case 1: rb_tile<VT, XT, AT, NV, VW, KC, 1, U>(acc, pk, s0, s1, sub, sX); break;
            case 2: if constexpr (KC >= 2) rb_tile<VT, XT, AT, NV, VW, KC, 2, U>(acc, pk, s0, s1, sub, sX); break;
            case 3: if constexpr (KC >= 3) rb_tile<VT, XT, AT, NV, VW, KC, 3, U>(acc, pk, s0, s1, sub, sX); break;
            default: if constexpr (KC >= 4) rb_tile<VT, XT, AT, NV, VW, KC, 4, U>(acc, pk, s0, s1, sub, sX); break;
        }
    }
#pragma unroll
    for (int k = 0; k < KC; k++) {
        if (k >= kn) break;
        const int c = ch0 + wv + W * k;
#pragma unroll
        for (int v = 0; v < NV; v++) {
            acc[k][v] += __shfl_xor(acc[k][v], 16);
            acc[k][v] += __shfl_xor(acc[k][v], 32);
        }
        const int p = c * RB_R + (lane & 15);
        if (sub == 0 && p < n) {
#pragma unroll
            for (int v = 0; v < NV; v++) slot[(size_t)(lp + p) * NV + v] = acc[k][v];
        }
    }
}

// k_spmm_rbmf: the block product on the matrix cores.  Same units, tiles,
// X staging and chunk assignment as k_spmm_rblk; per 16 x 8 tile step and
// chunk, for each 16-vector slice of X, two 16 x 16 x 4 products (columns 0-3
// and 4-7 of the step): A = the lane's two values as loaded, B = X[column
// l / 16 (+ 4)][vector l mod 16 of the slice] from LDS, read once for all of
// the wave's chunks.  Accumulation on the matrix cores in the precision of
// the values (FP32 for FP32 H, FP64 for FP64 H).
template <typename T> struct MfmaT;
template <> struct MfmaT<float> {
    typedef float D __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ D mma(float a, float b, D c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // the result row element v of lane l holds (its column is l mod 16)
    static __device__ __forceinline__ int row(int l, int v) { return 4 * (l >> 4) + v; }
};
template <> struct MfmaT<double> {
    typedef double D __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ D mma(double a, double b, D c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int l, int v) { return (l >> 4) + 4 * v; }
};

template <typename VT, int NV, int KC, int KN, int UN>
__device__ __forceinline__ void rbmf_steps(typename MfmaT<VT>::D (&d)[KC][(NV + 15) / 16],
                                           const typename VecN<VT, 2>::type *const (&pk)[KC], int s, int s0, int lane,
                                           const VT *sX) {
    constexpr int VW = sizeof(VT) == 4 ? 4 : 2;
    constexpr int NVP = (NV + VW - 1) / VW * VW;
    constexpr int NT = (NV + 15) / 16;
    typedef typename VecN<VT, 2>::type V2;
    V2 vv[KN][UN];
#pragma unroll
    for (int k = 0; k < KN; k++)
#pragma unroll
        for (int u = 0; u < UN; u++) vv[k][u] = __builtin_nontemporal_load(pk[k] + (size_t)(s + u) * 64);
    const int cv = lane & 15;
#pragma unroll
    for (int u = 0; u < UN; u++) {
        const VT *xr = sX + (8 * (s + u - s0) + (lane >> 4)) * NVP;  // column 8 (s+u) + l / 16 of the tile
#pragma unroll
        for (int h = 0; h < 2; h++) {
#pragma unroll
            for (int nt = 0; nt < NT; nt++) {
                const int v = 16 * nt + cv;
                const VT xb = xr[4 * h * NVP + (v < NV ? v : NV - 1)];
                const VT b = v < NV ? xb : (VT)0;
#pragma unroll
                for (int k = 0; k < KN; k++) d[k][nt] = MfmaT<VT>::mma(vv[k][u][h], b, d[k][nt]);
            }
        }
    }
}

template <typename VT, int NV, int KC, int KN, int U>
__device__ __forceinline__ void rbmf_tile(typename MfmaT<VT>::D (&d)[KC][(NV + 15) / 16],
                                          const typename VecN<VT, 2>::type *const (&pk)[KC], int s0, int s1, int lane,
                                          const VT *sX) {
    int s = s0;
    for (; s + U <= s1; s += U) rbmf_steps<VT, NV, KC, KN, U>(d, pk, s, s0, lane, sX);
    for (; s < s1; s++) rbmf_steps<VT, NV, KC, KN, 1>(d, pk, s, s0, lane, sX);
}

template <typename VT, int NV, int W, int KC, int U>
__global__ __launch_bounds__(64 * W) void k_spmm_rbmf(const int *__restrict__ unit, const int64_t *__restrict__ list_ptr,
                                                      const int *__restrict__ list_det,
                                                      const int64_t *__restrict__ boff, const VT *__restrict__ val,
                                                      int TC, const VT *__restrict__ X, int ldx,
                                                      VT *__restrict__ slot) {
    constexpr int VW = sizeof(VT) == 4 ? 4 : 2;
    constexpr int NVV = (NV + VW - 1) / VW;
    constexpr int NVP = NVV * VW;
    constexpr int NT = (NV + 15) / 16;
    static_assert(KC >= 1 && KC <= 4, "chunks per wave");
    typedef VT V __attribute__((ext_vector_type(VW)));
    typedef typename VecN<VT, 2>::type V2;
    typedef typename MfmaT<VT>::D D;
    extern __shared__ __align__(16) char rbm_lds[];
    VT *sX = (VT *)rbm_lds;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int lid = unit[3 * blockIdx.x], ch0 = unit[3 * blockIdx.x + 1], ch1 = unit[3 * blockIdx.x + 2];
    const int64_t lp = list_ptr[lid];
    const int n = (int)(list_ptr[lid + 1] - lp);
    const int steps = (n + 7) >> 3;
    const V2 *sv = (const V2 *)(val + boff[lid]);
    const int kn = ch1 - ch0 - wv <= 0 ? 0 : min(KC, (ch1 - ch0 - wv + W - 1) / W);
    const V2 *pk[KC];
#pragma unroll
    for (int k = 0; k < KC; k++) pk[k] = sv + (size_t)(ch0 + wv + W * min(k, max(kn - 1, 0))) * steps * 64 + lane;
    D d[KC][NT];
#pragma unroll
    for (int k = 0; k < KC; k++)
#pragma unroll
        for (int t = 0; t < NT; t++) d[k][t] = (D)(VT)0;
    for (int j0 = 0; j0 < n; j0 += TC) {
        const int nj = min(TC, n - j0);
        const int njp = (nj + 7) & ~7;
        __syncthreads();
        for (int e = tid; e < njp * NVV; e += 64 * W) {
            const int r = e / NVV, q = e - r * NVV;
            V x = (V)(VT)0;
            if (r < nj) x = *(const V *)(X + (size_t)list_det[lp + j0 + r] * ldx + q * VW);
            *(V *)(sX + r * NVP + q * VW) = x;
        }
        __syncthreads();
        const int s0 = j0 >> 3, s1 = s0 + (njp >> 3);
        switch (kn) {  // wave-uniform
            case 0: break;
            case 1: rbmf_tile<VT, NV, KC, 1, U>(d, pk, s0, s1, lane, sX); break;
            case 2: if constexpr (KC >= 2) rbmf_tile<VT, NV, KC, 2, U>(d, pk, s0, s1, lane, sX); break;
            case 3: if constexpr (KC >= 3) rbmf_tile<VT, NV, KC, 3, U>(d, pk, s0, s1, lane, sX); break;
            default: if constexpr (KC >= 4) rbmf_tile<VT, NV, KC, 4, U>(d, pk, s0, s1, lane, sX); break;
        }
    }
#pragma unroll
    for (int k = 0; k < KC; k++) {
        if (k >= kn) break;
        const int c = ch0 + wv + W * k;
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int v = 16 * t + (lane & 15);
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int p = c * RB_R + MfmaT<VT>::row(lane, e);
                if (p < n && v < NV) slot[(size_t)(lp + p) * NV + v] = d[k][t][e];
            }
        }
    }
}

// slot_idx[i * nres + q] = list slot of determinant i in its q-th list
__global__ void k_rb_slot_idx(int nd, ResView R, const int *__restrict__ det_pos, int *__restrict__ slot_idx) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nd * R.nres) return;
    slot_idx[e] = (int)(R.list_ptr[R.det_res[e]] + det_pos[e]);
}

// Y[i, :] = sum over the determinant's lists (in list order) of its slot,
// plus shift X[i, :]: one thread per determinant
template <typename XT, typename AT, int NV>
__global__ __launch_bounds__(256) void k_rb_combine(int nd, int nres, const int *__restrict__ slot_idx,
                                                    const AT *__restrict__ slot, const XT *__restrict__ X, int ldx,
                                                    XT *__restrict__ Y, int ldy, AT shift) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nd) return;
    AT s[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) s[v] = 0;
    for (int q = 0; q < nres; q++) {
        const AT *sl = slot + (size_t)slot_idx[(size_t)i * nres + q] * NV;
#pragma unroll
        for (int v = 0; v < NV; v++) s[v] += sl[v];
    }
#pragma unroll
    for (int v = 0; v < NV; v++) Y[(size_t)i * ldy + v] = (XT)(s[v] + shift * (AT)X[(size_t)i * ldx + v]);
}

// Wide-panel variant for large Nd, where 1024-column LDS panels would cut
// each row into runs of a few elements: the same sliced layout with panels
// of 65536 columns (16-bit columns still), X gathered straight from memory.
// A tile is one row block x one panel, tiles run panel by panel (blockIdx
// panel-major), so the workgroups resident at any time share one X panel
// (65536 rows, 2 MB for 8 FP32 columns) and every XCD's L2 holds it while
// the non-temporal matrix stream passes by.  No LDS, no barriers; per-panel
// partial sums are added in panel order by k_spmm_panel_sum.
constexpr int SELLG_P = 65536;

template <typename VT, typename XT, typename AT, int NV, int VW, int W, int U>
__global__ __launch_bounds__(64 * W) void k_spmm_sellg(int nd, int npan, int nrb, const int64_t *__restrict__ off,
                                                       const uint32_t *__restrict__ scol,
                                                       const VT *__restrict__ sval, const XT *__restrict__ X,
                                                       int ldx, XT *__restrict__ Y, int ldy, AT *__restrict__ part,
                                                       AT shift) {
    constexpr int NVV = (NV + VW - 1) / VW;
    constexpr int NVP = NVV * VW;
    typedef XT V __attribute__((ext_vector_type(VW)));
    typedef typename VecN<VT, 2>::type V2;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int p = blockIdx.x / nrb, rb = blockIdx.x - p * nrb;
    const int c = rb * W + wv;
    if (c >= (nd + SELL_R - 1) / SELL_R) return;  // wave-uniform, no barriers follow
    const XT *Xp = X + (size_t)p * SELLG_P * ldx;
    const V2 *sv = (const V2 *)sval;
    AT acc[NVP];
#pragma unroll
    for (int v = 0; v < NVP; v++) acc[v] = 0;
    const int64_t seg = (int64_t)c * npan + p;
    const int64_t e = off[seg + 1];
    for (int64_t b = off[seg]; b < e; b += U) {
        uint32_t cc[U];
        V2 vv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (b + u < e) {
                cc[u] = __builtin_nontemporal_load(scol + (b + u) * 64 + lane);
                vv[u] = __builtin_nontemporal_load(sv + (b + u) * 64 + lane);
            } else {
                cc[u] = 0;
                vv[u] = (V2)(VT)0;
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int cl = h ? (int)(cc[u] >> 16) : (int)(cc[u] & 0xffffu);
                const AT a = (AT)vv[u][h];
                if (a != (AT)0) {  // padding slots hold 0 and read no X
                    const XT *xr = Xp + (size_t)cl * ldx;
#pragma unroll
                    for (int q = 0; q < NVV; q++) {
                        const V x = *(const V *)(xr + q * VW);
#pragma unroll
                        for (int v = 0; v < VW; v++) acc[q * VW + v] += a * (AT)x[v];
                    }
                }
            }
        }
    }
#pragma unroll
    for (int v = 0; v < NV; v++) {
        acc[v] += __shfl_xor(acc[v], 1);
        acc[v] += __shfl_xor(acc[v], 2);
    }
    const int row = c * SELL_R + (lane >> 2);
    if ((lane & 3) == 0 && row < nd) {
        if (part) {
#pragma unroll
            for (int v = 0; v < NV; v++) part[((size_t)p * nd + row) * NV + v] = acc[v];
        } else {
#pragma unroll
            for (int v = 0; v < NV; v++) Y[(size_t)row * ldy + v] = (XT)(acc[v] + shift * (AT)X[(size_t)row * ldx + v]);
        }
    }
}

// ---------------------------------------------------------------- dispatch
enum SpmmVariant {
    SV_SCALAR = 0, SV_GROUP = 1, SV_LANE1 = 2, SV_LANE2 = 3, SV_LANE1_W16 = 4, SV_LANE2_W16 = 5,
    SV_PANEL2 = 6, SV_PANEL4 = 7, SV_PANEL8 = 8, SV_XLDS2 = 9, SV_XLDS4 = 10, SV_XG4x2 = 11, SV_XG4x4 = 12,
    SV_XG2x4 = 13, SV_XLDS2U2 = 14, SV_XLDS1U2 = 15, SV_XP2 = 16, SV_XP1 = 17, SV_SELL1 = 18, SV_SELL2 = 19,
    SV_SELL3 = 20, SV_SELLG1 = 21, SV_SELLG2 = 22, SV_SELLQ1 = 23, SV_SELLQ2 = 24, SV_SELLQ3 = 25,
    SV_RBLK1 = 26, SV_RBLK2 = 27, SV_RBMF1 = 28, SV_RBMF2 = 29, SV_COUNT = 30
};
// the residual-block shapes: SV_RBLK1 ... SV_RB_LAST
constexpr int SV_RB_LAST = SV_RBMF2;
static const char *const kVariantName[SV_COUNT] = {
    "k_spmm", "k_spmm_vec", "k_spmm_lane<U=1,4 rows/block>", "k_spmm_lane<U=2,4 rows/block>",
    "k_spmm_lane<U=1,16 rows/block>", "k_spmm_lane<U=2,16 rows/block>", "k_spmm_panel<K=2>", "k_spmm_panel<K=4>",
    "k_spmm_panel<K=8>", "k_spmm_xlds<2 rows/wave>", "k_spmm_xlds<4 rows/wave>",
    "k_spmm_xldsg<4 rows x 2/wave>", "k_spmm_xldsg<4 rows x 4/wave>", "k_spmm_xldsg<2 rows x 4/wave>",
    "k_spmm_xlds<2 rows/wave,U=2>", "k_spmm_xlds<1 row/wave,U=2>", "k_spmm_xlds2<2 rows/wave>",
    "k_spmm_xlds2<1 row/wave>", "k_spmm_sell<16 waves>", "k_spmm_sell<8 waves>",
    "k_spmm_sell<16 waves,4|2 chunks/wave>", "k_spmm_sellg<4 waves>", "k_spmm_sellg<16 waves>",
    "k_spmm_sellq<16 waves,4|2 chunks,2 steps>", "k_spmm_sellq<8 waves,4|2 chunks,2 steps,prefetch>",
    "k_spmm_sellq<16 waves,2|1 chunks,4 steps,prefetch>", "k_spmm_rblk<8 waves,4 chunks>",
    "k_spmm_rblk<16 waves,2 chunks>", "k_spmm_rbmf<8 waves,4 chunks>", "k_spmm_rbmf<16 waves,2 chunks>"};
static int panel_count(int variant) {
    return variant >= SV_PANEL2 && variant <= SV_PANEL8 ? 2 << (variant - SV_PANEL2) : 1;
}

template <typename VT, typename XT, typename AT, int NV>
static void launch_nv(int variant, unsigned grid, hipStream_t s, int nd, const int64_t *rp, const int *col,
                      const VT *val, const XT *X, int ldx, XT *Y, int ldy, AT shift, const int64_t *pp, AT *part) {
    constexpr int VW = sizeof(XT) == 4 ? 4 : 2;
    const unsigned grid16 = (unsigned)(((nd + 15) / 16 + 7) / 8 * 8);
    if (variant == SV_XLDS2) {
        k_spmm_xlds<VT, XT, AT, NV, VW, 2, 1><<<(unsigned)((nd + 31) / 32), 1024, 0, s>>>(nd, rp, col, val, X, ldx, Y,
                                                                                      ldy, shift);
        return;
    }
    if (variant == SV_XP2) {
        k_spmm_xlds2<VT, XT, AT, NV, VW, 2><<<(unsigned)((nd + 31) / 32), 1024, 0, s>>>(nd, rp, col, val, X, ldx, Y,
                                                                                     ldy, shift);
        return;
    }
    if (variant == SV_XP1) {
        k_spmm_xlds2<VT, XT, AT, NV, VW, 1><<<(unsigned)((nd + 15) / 16), 1024, 0, s>>>(nd, rp, col, val, X, ldx, Y,
                                                                                     ldy, shift);
        return;
    }
    if (variant == SV_XLDS2U2) {
        k_spmm_xlds<VT, XT, AT, NV, VW, 2, 2><<<(unsigned)((nd + 31) / 32), 1024, 0, s>>>(nd, rp, col, val, X, ldx, Y,
                                                                                      ldy, shift);
        return;
    }
    if (variant == SV_XLDS1U2) {
        k_spmm_xlds<VT, XT, AT, NV, VW, 1, 2><<<(unsigned)((nd + 15) / 16), 1024, 0, s>>>(nd, rp, col, val, X, ldx, Y,
                                                                                      ldy, shift);
        return;
    }
    if (variant >= SV_XG4x2 && variant <= SV_XG2x4) {
        constexpr bool fit4 = NV <= 16, fit2 = NV <= 32;
        if (variant == SV_XG4x2 && fit4)
            k_spmm_xldsg<VT, XT, AT, NV, VW, (fit4 ? 4 : 1), 2><<<(unsigned)((nd + 127) / 128), 1024, 0, s>>>(
                nd, rp, col, val, X, ldx, Y, ldy, shift);
        else if (variant == SV_XG4x4 && fit4)
            k_spmm_xldsg<VT, XT, AT, NV, VW, (fit4 ? 4 : 1), 4><<<(unsigned)((nd + 255) / 256), 1024, 0, s>>>(
                nd, rp, col, val, X, ldx, Y, ldy, shift);
        else if (variant == SV_XG2x4 && fit2)
            k_spmm_xldsg<VT, XT, AT, NV, VW, (fit2 ? 2 : 1), 4><<<(unsigned)((nd + 127) / 128), 1024, 0, s>>>(
                nd, rp, col, val, X, ldx, Y, ldy, shift);
        else
            k_spmm_xlds<VT, XT, AT, NV, VW, 2, 1><<<(unsigned)((nd + 31) / 32), 1024, 0, s>>>(nd, rp, col, val, X, ldx,
                                                                                          Y, ldy, shift);
        return;
    }
    if (variant == SV_XLDS4) {
        k_spmm_xlds<VT, XT, AT, NV, VW, 4, 1><<<(unsigned)((nd + 63) / 64), 1024, 0, s>>>(nd, rp, col, val, X, ldx, Y,
                                                                                      ldy, shift);
        return;
    }
    if (variant >= SV_PANEL2) {
        const int K = panel_count(variant);
        const unsigned g = (unsigned)(K * ((nd + 3) / 4));
        k_spmm_panel<VT, XT, AT, NV, VW, 2><<<g, 256, 0, s>>>(nd, pp, col, val, X, ldx, part);
        const int64_t n = (int64_t)nd * NV;
        k_spmm_panel_sum<XT, AT><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(nd, NV, K, part, X, ldx, Y, ldy, shift);
        return;
    }
    switch (variant) {
        case SV_GROUP: k_spmm_vec<VT, XT, AT, NV, VW><<<grid, 256, 0, s>>>(nd, rp, col, val, X, ldx, Y, ldy, shift); break;
        case SV_LANE1: k_spmm_lane<VT, XT, AT, NV, VW, 1, 4><<<grid, 256, 0, s>>>(nd, rp, col, val, X, ldx, Y, ldy, shift); break;
        case SV_LANE2: k_spmm_lane<VT, XT, AT, NV, VW, 2, 4><<<grid, 256, 0, s>>>(nd, rp, col, val, X, ldx, Y, ldy, shift); break;
        case SV_LANE1_W16: k_spmm_lane<VT, XT, AT, NV, VW, 1, 16><<<grid16, 1024, 0, s>>>(nd, rp, col, val, X, ldx, Y, ldy, shift); break;
        case SV_LANE2_W16: k_spmm_lane<VT, XT, AT, NV, VW, 2, 16><<<grid16, 1024, 0, s>>>(nd, rp, col, val, X, ldx, Y, ldy, shift); break;
        default: k_spmm<VT, XT, AT, NV><<<grid, 256, 0, s>>>(nd, rp, col, val, X, ldx, Y, ldy, shift); break;
    }
}

// Scratch of one launch (partial sums): the matrix's own buffer, grown on
// demand, or in an allocation-free call the caller-independent file-scope
// array, checked for size before anything is launched.
static char *launch_scratch(Matrix &M, DBuf<char> &own, size_t need, hipStream_t s) {
    if (M.noalloc) {
        PCI_REQUIRE(need <= M.ext_cap, PCI_E_ARG,
                    "pci_gpu_spmm_noalloc: this product needs " + std::to_string(need) +
                        " bytes of scratch, the entry point has " + std::to_string(M.ext_cap));
        return M.ext;
    }
    if (own.n < need) {
        PCI_HIP(hipStreamSynchronize(s));
        own.alloc(need);
    }
    return own.p;
}

// In an allocation-free call a copy the kernel needs must already exist.
static void require_built(const Matrix &M, bool built) {
    PCI_REQUIRE(!M.noalloc || built, PCI_E_ARG,
                "pci_gpu_spmm_noalloc: the chosen kernel's copy of H is not built; call pci_gpu_spmm (or a "
                "Davidson solve) with this block width and precision first");
}

// One launch of k_spmm_sell: tiles of W chunks x pg panels, pg chosen
// so that there are about `target` tiles (several per CU); more than one
// panel group writes partial sums that k_spmm_panel_sum adds in order.
template <typename VT, typename XT, typename AT, int NV, int VW, int W, int KC, int U, bool SKIP>
static void sell_go(Matrix &M, const SellCopy &S, const XT *X, int ldx, XT *Y, int ldy, AT shift, int target,
                    hipStream_t s) {
    const int nd = M.nd;
    const int nrb = (int)((S.nchunk + W * KC - 1) / (W * KC));
    int ng = std::max(1, std::min(S.npan, (target + nrb - 1) / nrb));
    const int pg = (S.npan + ng - 1) / ng;
    ng = (S.npan + pg - 1) / pg;
    AT *part = nullptr;
    if (ng > 1) part = (AT *)launch_scratch(M, M.part, (size_t)ng * nd * NV * sizeof(AT), s);
    const size_t shm = (size_t)S.P * sell_ldx<XT>(NV) * sizeof(XT);
    k_spmm_sell<VT, XT, AT, NV, VW, W, KC, U, SKIP><<<(unsigned)(nrb * ng), 64 * W, shm, s>>>(
        nd, S.P, S.npan, pg, nrb, S.off.p, S.col.p, (const VT *)S.val.p, X, ldx, Y, ldy, part, shift);
    if (part) {
        const int64_t n = (int64_t)nd * NV;
        k_spmm_panel_sum<XT, AT><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(nd, NV, ng, part, X, ldx, Y, ldy,
                                                                            shift);
    }
}

// One launch of k_spmm_sellq over the same tiles as sell_go.
template <typename VT, typename XT, typename AT, int NV, int VW, int W, int KC, int U, bool PF>
static void sellq_go(Matrix &M, const SellCopy &S, const XT *X, int ldx, XT *Y, int ldy, AT shift, int target,
                     hipStream_t s) {
    const int nd = M.nd;
    const int nrb = (int)((S.nchunk + W * KC - 1) / (W * KC));
    int ng = std::max(1, std::min(S.npan, (target + nrb - 1) / nrb));
    const int pg = (S.npan + ng - 1) / ng;
    ng = (S.npan + pg - 1) / pg;
    AT *part = nullptr;
    if (ng > 1) part = (AT *)launch_scratch(M, M.part, (size_t)ng * nd * NV * sizeof(AT), s);
    const size_t shm = (size_t)S.P * sell_ldx<XT>(NV) * sizeof(XT);
    k_spmm_sellq<VT, XT, AT, NV, VW, W, KC, U, PF><<<(unsigned)(nrb * ng), 64 * W, shm, s>>>(
        nd, S.P, S.npan, pg, nrb, S.off.p, S.col.p, (const VT *)S.val.p, X, ldx, Y, ldy, part, shift);
    if (part) {
        const int64_t n = (int64_t)nd * NV;
        k_spmm_panel_sum<XT, AT><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(nd, NV, ng, part, X, ldx, Y, ldy,
                                                                            shift);
    }
}

template <typename VT, typename XT, typename AT, int NV>
static void launch_sell(Matrix &M, int variant, const VT *val, const XT *X, int ldx, XT *Y, int ldy, AT shift,
                        hipStream_t s) {
    constexpr int VW = sizeof(XT) == 4 ? 4 : 2;
    constexpr int NVP = (NV + VW - 1) / VW * VW;
    SellCopy &S = M.sell[sizeof(VT) == 4 ? 0 : 1];
    if constexpr (NVP <= 24) {
        // a copy built for a narrower block has panels too wide for this
        // width's LDS panel: rebuild it for this width (its panels then also
        // serve every narrower width)
        if (S.built && (int64_t)S.P * sell_ldx<XT>(NV) * (int64_t)sizeof(XT) > SELL_LDS) {
            require_built(M, false);
            PCI_HIP(hipStreamSynchronize(s));
            S.release();
        }
        if (!S.built && S.eligible) {
            require_built(M, false);
            DBuf<char> work;
            sell_build<VT>(M, S, val, sell_panel<XT>(NV), s, work);
        }
        if (S.built) {
            // SELL3: several chunks per wavefront (1024 FP32 / 512 FP64 rows per
            // workgroup), for matrices whose rows have few elements per panel,
            // where one X panel copy must serve more rows; above 16 padded
            // columns its accumulators spill, so it runs one chunk per wave
            constexpr int KC3 = NVP > 16 ? 1 : (sizeof(XT) == 4 ? 4 : 2);
            // SELL1's target tile count per precision: half as many panel
            // groups (fewer partial sums, longer tiles) measured faster in FP64
            // on C2 (1.10 -> 1.04 ms at 10 columns) and slower in FP32
            constexpr int T1 = sizeof(XT) == 4 ? 4096 : 2048;
            if constexpr (NVP <= 16) {
                constexpr int KQ = sizeof(XT) == 4 ? 4 : 2;
                if (variant == SV_SELLQ1) {
                    sellq_go<VT, XT, AT, NV, VW, 16, KQ, 2, false>(M, S, X, ldx, Y, ldy, shift, 4096, s);
                    return;
                }
                if (variant == SV_SELLQ2) {
                    sellq_go<VT, XT, AT, NV, VW, 8, KQ, 2, true>(M, S, X, ldx, Y, ldy, shift, 4096, s);
                    return;
                }
                if (variant == SV_SELLQ3) {
                    sellq_go<VT, XT, AT, NV, VW, 16, KQ / 2, 4, true>(M, S, X, ldx, Y, ldy, shift, 4096, s);
                    return;
                }
            }
            if (variant == SV_SELL1) sell_go<VT, XT, AT, NV, VW, 16, 1, 4, true>(M, S, X, ldx, Y, ldy, shift, T1, s);
            else if (variant == SV_SELL2)
                sell_go<VT, XT, AT, NV, VW, 8, 1, 4, true>(M, S, X, ldx, Y, ldy, shift, 8192, s);
            else sell_go<VT, XT, AT, NV, VW, 16, KC3, 4, true>(M, S, X, ldx, Y, ldy, shift, 4096, s);
            return;
        }
    }
    PCI_REQUIRE(!M.blocks_only, PCI_E_UNSUPPORTED, "H held as residual blocks only: block widths up to 24");
    const int nd = M.nd;
    k_spmm_xlds<VT, XT, AT, NV, VW, 2, 1><<<(unsigned)((nd + 31) / 32), 1024, 0, s>>>(nd, M.rp.p, M.col.p, val, X,
                                                                                  ldx, Y, ldy, shift);
}

template <typename VT, typename XT, typename AT, int NV, int W, int U>
static void sellg_go(Matrix &M, const SellCopy &S, const XT *X, int ldx, XT *Y, int ldy, AT shift, hipStream_t s) {
    constexpr int VW = sizeof(XT) == 4 ? 4 : 2;
    const int nd = M.nd;
    const int nrb = (int)((S.nchunk + W - 1) / W);
    AT *part = nullptr;
    if (S.npan > 1) part = (AT *)launch_scratch(M, M.part, (size_t)S.npan * nd * NV * sizeof(AT), s);
    k_spmm_sellg<VT, XT, AT, NV, VW, W, U><<<(unsigned)((int64_t)nrb * S.npan), 64 * W, 0, s>>>(
        nd, S.npan, nrb, S.off.p, S.col.p, (const VT *)S.val.p, X, ldx, Y, ldy, part, shift);
    if (part) {
        const int64_t n = (int64_t)nd * NV;
        k_spmm_panel_sum<XT, AT><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(nd, NV, S.npan, part, X, ldx, Y, ldy,
                                                                            shift);
    }
}

template <typename VT, typename XT, typename AT, int NV>
static void launch_sellg(Matrix &M, int variant, const VT *val, const XT *X, int ldx, XT *Y, int ldy, AT shift,
                         hipStream_t s) {
    constexpr int VW = sizeof(XT) == 4 ? 4 : 2;
    constexpr int NVP = (NV + VW - 1) / VW * VW;
    SellCopy &S = M.sellg[sizeof(VT) == 4 ? 0 : 1];
    if constexpr (NVP <= 24) {
        if (!S.built && S.eligible) {
            require_built(M, false);
            DBuf<char> work;
            sell_build<VT>(M, S, val, SELLG_P, s, work);
        }
        if (S.built) {
            if (variant == SV_SELLG1) sellg_go<VT, XT, AT, NV, 4, 4>(M, S, X, ldx, Y, ldy, shift, s);
            else sellg_go<VT, XT, AT, NV, 16, 4>(M, S, X, ldx, Y, ldy, shift, s);
            return;
        }
    }
    PCI_REQUIRE(!M.blocks_only, PCI_E_UNSUPPORTED, "H held as residual blocks only: block widths up to 24");
    const int nd = M.nd;
    k_spmm_xlds<VT, XT, AT, NV, VW, 2, 1><<<(unsigned)((nd + 31) / 32), 1024, 0, s>>>(nd, M.rp.p, M.col.p, val, X,
                                                                                  ldx, Y, ldy, shift);
}

// block widths with vector kernels instantiated (others use k_spmm)
static bool vec_width(int n) { return (n >= 1 && n <= 24) || n == 28 || n == 32; }

// ------------------------------------------------- residual-block copy
constexpr int RB_CPU = 32;  // row chunks per unit (W x KC of both launch shapes)

template <typename VT>
static void rblk_fill_launch(int W, int nd, const ResView &R, const int *combo, const int *det_pos, const Matrix &M,
                             const VT *val, const int64_t *boff, VT *out, int *bad, hipStream_t s) {
    const unsigned g = (unsigned)((nd + 3) / 4);
    switch (W) {
#define PCI_W(w) case w: k_rb_fill<VT, w><<<g, 256, 0, s>>>(nd, R, combo, det_pos, M.rp.p, M.col.p, val, boff, out, bad); break;
        PCI_W(1) PCI_W(2) PCI_W(3) PCI_W(4) PCI_W(5) PCI_W(6) PCI_W(7) PCI_W(8)
        PCI_W(9) PCI_W(10) PCI_W(11) PCI_W(12) PCI_W(13) PCI_W(14) PCI_W(15) PCI_W(16)
#undef PCI_W
        default: throw Error(PCI_E_UNSUPPORTED, "residual blocks: more than 16 occupation words");
    }
    PCI_HIP(hipGetLastError());
}

// Block offsets and units of M's lists: each list's chunks split into units
// of at most RB_CPU chunks, units ordered largest list first (so the long
// ones start early).  Returns nvals; B.boff and B.unit are uploaded and
// B.nunit, B.maxn set (B.nvals is set once the values exist).
int64_t rblk_layout(Matrix &M, RblkCopy &B, hipStream_t s) {
    const ResView &R = M.res;
    const int64_t nl = R.nlists;
    std::vector<int64_t> lp(nl + 1);
    PCI_HIP(hipMemcpyAsync(lp.data(), R.list_ptr, (nl + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    PCI_HIP(hipStreamSynchronize(s));
    std::vector<int64_t> boff(nl + 1, 0);
    struct U3 { int n, lid, c0, c1; };
    std::vector<U3> units;
    int maxn = 0;
    for (int64_t r = 0; r < nl; r++) {
        const int n = (int)(lp[r + 1] - lp[r]);
        const int chunks = (n + RB_R - 1) / RB_R, steps = (n + 7) / 8;
        boff[r + 1] = boff[r] + (int64_t)chunks * steps * 128;
        maxn = std::max(maxn, n);
        const int nu = (chunks + RB_CPU - 1) / RB_CPU;
        for (int u = 0; u < nu; u++)
            units.push_back({n, (int)r, (int)((int64_t)chunks * u / nu), (int)((int64_t)chunks * (u + 1) / nu)});
    }
    std::stable_sort(units.begin(), units.end(), [](const U3 &a, const U3 &b) { return a.n > b.n; });
    std::vector<int> ut(3 * units.size());
    for (size_t u = 0; u < units.size(); u++) {
        ut[3 * u] = units[u].lid;
        ut[3 * u + 1] = units[u].c0;
        ut[3 * u + 2] = units[u].c1;
    }
    upload(B.unit, ut.data(), ut.size(), s);
    upload(B.boff, boff.data(), boff.size(), s);
    B.maxn = maxn;
    B.nunit = (int)units.size();
    return boff[nl];
}

// Each determinant's position in each of its lists and the matching list
// slot (once per matrix; shared by both precisions).
void rblk_positions(Matrix &M, hipStream_t s) {
    if (M.det_pos.p) return;
    const ResView &R = M.res;
    PCI_REQUIRE((int64_t)M.nd * R.nres < ((int64_t)1 << 31), PCI_E_UNSUPPORTED, "more than 2^31 list entries");
    M.det_pos.alloc((size_t)M.nd * R.nres);
    k_rb_pos<<<(unsigned)R.nlists, 64, 0, s>>>(R, M.det_pos.p);
    PCI_HIP(hipGetLastError());
    M.slot_idx.alloc((size_t)M.nd * R.nres);
    const int64_t ne = (int64_t)M.nd * R.nres;
    k_rb_slot_idx<<<(unsigned)((ne + 255) / 256), 256, 0, s>>>(M.nd, R, M.det_pos.p, M.slot_idx.p);
    PCI_HIP(hipGetLastError());
}

// combo[mask] = index of the rsize-subset `mask` of the Ne positions in the
// order of a determinant's residual lists (lexicographic; -1 for masks of
// another size)
static std::vector<int> rb_combos(int ne, int rsize) {
    std::vector<int> combo(1 << ne, -1);
    std::vector<int> idx(rsize);
    for (int i = 0; i < rsize; i++) idx[i] = i;
    int q = 0;
    while (true) {
        int m = 0;
        for (int i = 0; i < rsize; i++) m |= 1 << idx[i];
        combo[m] = q++;
        int i = rsize - 1;
        while (i >= 0 && idx[i] == ne - rsize + i) i--;
        if (i < 0) break;
        idx[i]++;
        for (int j = i + 1; j < rsize; j++) idx[j] = idx[j - 1] + 1;
    }
    return combo;
}

// Builds B from the resident full CSR (values val: FP64, or FP32 with the
// shifted diagonal), or marks it ineligible: no residual lists for this
// matrix, blocks holding over 1.5x the stored elements (lists much longer
// than the rows they couple), not enough device memory (2 GB margin), an
// element outside the lists (a CSR read from file that does not belong to
// this space), or PCI_RBLK_OFF=1.
template <typename VT>
static void rblk_build(Matrix &M, RblkCopy &B, const VT *val, hipStream_t s) {
    const char *off = getenv("PCI_RBLK_OFF");
    if (!M.has_res || (off && atoi(off) != 0) || M.res.ne > RB_MAXNE) {
        B.eligible = false;
        return;
    }
    const ResView &R = M.res;
    try {
        const int64_t nvals = rblk_layout(M, B, s);
        if ((double)nvals > 1.5 * (double)M.nnz_full + 1e6) {
            B.release();
            B.eligible = false;
            return;
        }
        size_t fr = 0, tot = 0;
        PCI_HIP(hipMemGetInfo(&fr, &tot));
        const size_t need = (size_t)nvals * sizeof(VT) + (M.det_pos.p ? 0 : (size_t)M.nd * R.nres * 2 * sizeof(int));
        if (need + ((size_t)2 << 30) > fr)
            throw Error(PCI_E_NOMEM, "residual-block copy of " + std::to_string(need) + " bytes does not fit");
        rblk_positions(M, s);
        const std::vector<int> combo = rb_combos(R.ne, R.rsize);
        DBuf<int> dcombo, bad;
        upload(dcombo, combo.data(), combo.size(), s);
        bad.alloc(1);
        PCI_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), s));
        B.val.alloc((size_t)nvals * sizeof(VT));
        PCI_HIP(hipMemsetAsync(B.val.p, 0, B.val.bytes(), s));
        rblk_fill_launch<VT>(R.W, M.nd, R, dcombo.p, M.det_pos.p, M, val, B.boff.p, (VT *)B.val.p, bad.p, s);
        int hbad = 0;
        PCI_HIP(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
        PCI_HIP(hipStreamSynchronize(s));
        if (hbad) {
            B.release();
            B.eligible = false;
            return;
        }
        B.nvals = nvals;
    } catch (const Error &e) {
        if (e.code != PCI_E_NOMEM) throw;
        (void)hipGetLastError();
        (void)hipStreamSynchronize(s);
        B.release();
        B.eligible = false;
        return;
    }
    B.built = true;
}

template <typename VT, typename XT, typename AT, int NV, int W, int KC, int U, bool MF>
static void rblk_go(Matrix &M, const RblkCopy &B, const XT *X, int ldx, XT *Y, int ldy, AT shift, hipStream_t s) {
    constexpr int VW = sizeof(XT) == 4 ? 4 : 2;
    constexpr int NVP = (NV + VW - 1) / VW * VW;
    static_assert(W * KC == RB_CPU, "unit size");
    int TC = (65536 / (NVP * (int)sizeof(XT))) & ~7;
    TC = std::min(TC, (B.maxn + 7) & ~7);
    const size_t shm = (size_t)TC * NVP * sizeof(XT);
    AT *slot = (AT *)launch_scratch(M, M.slot, (size_t)M.nd * M.res.nres * NV * sizeof(AT), s);
    if constexpr (MF) {
        static_assert(std::is_same<VT, XT>::value && std::is_same<VT, AT>::value, "one precision on the matrix cores");
        k_spmm_rbmf<VT, NV, W, KC, U><<<(unsigned)B.nunit, 64 * W, shm, s>>>(
            B.unit.p, M.res.list_ptr, M.res.list_det, B.boff.p, (const VT *)B.val.p, TC, X, ldx, slot);
    } else {
        k_spmm_rblk<VT, XT, AT, NV, VW, W, KC, U><<<(unsigned)B.nunit, 64 * W, shm, s>>>(
            B.unit.p, M.res.list_ptr, M.res.list_det, B.boff.p, (const VT *)B.val.p, TC, X, ldx, slot);
    }
    k_rb_combine<XT, AT, NV><<<(unsigned)((M.nd + 255) / 256), 256, 0, s>>>(M.nd, M.res.nres, M.slot_idx.p, slot, X,
                                                                           ldx, Y, ldy, shift);
}

template <typename VT, typename XT, typename AT, int NV>
static void launch_rblk(Matrix &M, int variant, const VT *val, const XT *X, int ldx, XT *Y, int ldy, AT shift,
                        hipStream_t s) {
    constexpr int VW = sizeof(XT) == 4 ? 4 : 2;
    constexpr int NVP = (NV + VW - 1) / VW * VW;
    RblkCopy &B = M.rblk[sizeof(VT) == 4 ? 0 : 1];
    if constexpr (NVP <= 24) {
        if (!B.built && B.eligible) {
            require_built(M, false);
            rblk_build<VT>(M, B, val, s);
        }
        if (B.built) {
            // matrix loads in flight per chunk: fewer for wide FP64 blocks, whose
            // sums already take most of the registers
            constexpr int U = sizeof(AT) == 8 && NV > 12 ? 2 : 4;
            switch (variant) {
                case SV_RBLK1: rblk_go<VT, XT, AT, NV, 8, 4, U, false>(M, B, X, ldx, Y, ldy, shift, s); break;
                case SV_RBLK2: rblk_go<VT, XT, AT, NV, 16, 2, U, false>(M, B, X, ldx, Y, ldy, shift, s); break;
                case SV_RBMF1: rblk_go<VT, XT, AT, NV, 8, 4, 4, true>(M, B, X, ldx, Y, ldy, shift, s); break;
                default: rblk_go<VT, XT, AT, NV, 16, 2, 4, true>(M, B, X, ldx, Y, ldy, shift, s); break;
            }
            return;
        }
    }
    PCI_REQUIRE(!M.blocks_only, PCI_E_UNSUPPORTED, "H held as residual blocks only: block widths up to 24");
    const int nd = M.nd;
    k_spmm_xlds<VT, XT, AT, NV, VW, 2, 1><<<(unsigned)((nd + 31) / 32), 1024, 0, s>>>(nd, M.rp.p, M.col.p, val, X,
                                                                                  ldx, Y, ldy, shift);
}

// panel offsets and partial-sum scratch for K-panel variants (built once per matrix)
template <typename AT>
static void panel_setup(Matrix &M, int variant, int nvec, hipStream_t s, const int64_t **pp, AT **part) {
    *pp = nullptr;
    *part = nullptr;
    if (variant < SV_PANEL2 || variant > SV_PANEL8) return;
    const int K = panel_count(variant);
    DBuf<int64_t> &b = M.pan[variant - SV_PANEL2];
    if (!b.p) {
        require_built(M, false);
        b.alloc((size_t)(K + 1) * M.nd);
        k_panel_ptr<<<(M.nd + 255) / 256, 256, 0, s>>>(M.nd, K, M.rp.p, M.col.p, b.p);
        PCI_HIP(hipGetLastError());
    }
    *part = (AT *)launch_scratch(M, M.part, (size_t)K * M.nd * nvec * sizeof(AT), s);
    *pp = b.p;
}

template <typename VT, typename XT, typename AT>
static void launch_variant(Matrix &M, int variant, const VT *val, const XT *X, int ldx, XT *Y, int ldy, int nvec,
                           AT shift, hipStream_t s) {
    const int nd = M.nd;
    const int64_t *rp = M.rp.p;
    const int *col = M.col.p;
    const unsigned grid = (unsigned)(((nd + 3) / 4 + 7) / 8 * 8);
    const int64_t *pp;
    AT *part;
    panel_setup<AT>(M, variant, nvec, s, &pp, &part);
    if (variant >= SV_RBLK1 && variant <= SV_RB_LAST) {
        switch (nvec) {
#define PCI_NV(n) case n: launch_rblk<VT, XT, AT, n>(M, variant, val, X, ldx, Y, ldy, shift, s); break;
            PCI_NV(1) PCI_NV(2) PCI_NV(3) PCI_NV(4) PCI_NV(5) PCI_NV(6) PCI_NV(7) PCI_NV(8)
            PCI_NV(9) PCI_NV(10) PCI_NV(11) PCI_NV(12) PCI_NV(13) PCI_NV(14) PCI_NV(15) PCI_NV(16)
            PCI_NV(17) PCI_NV(18) PCI_NV(19) PCI_NV(20) PCI_NV(21) PCI_NV(22) PCI_NV(23) PCI_NV(24)
            PCI_NV(28) PCI_NV(32)
#undef PCI_NV
            default: throw Error(PCI_E_UNSUPPORTED, "SpMM block width not supported");
        }
        PCI_HIP(hipGetLastError());
        return;
    }
    if (variant >= SV_SELLG1 && variant <= SV_SELLG2) {
        switch (nvec) {
#define PCI_NV(n) case n: launch_sellg<VT, XT, AT, n>(M, variant, val, X, ldx, Y, ldy, shift, s); break;
            PCI_NV(1) PCI_NV(2) PCI_NV(3) PCI_NV(4) PCI_NV(5) PCI_NV(6) PCI_NV(7) PCI_NV(8)
            PCI_NV(9) PCI_NV(10) PCI_NV(11) PCI_NV(12) PCI_NV(13) PCI_NV(14) PCI_NV(15) PCI_NV(16)
            PCI_NV(17) PCI_NV(18) PCI_NV(19) PCI_NV(20) PCI_NV(21) PCI_NV(22) PCI_NV(23) PCI_NV(24)
            PCI_NV(28) PCI_NV(32)
#undef PCI_NV
            default: throw Error(PCI_E_UNSUPPORTED, "SpMM block width not supported");
        }
        PCI_HIP(hipGetLastError());
        return;
    }
    if ((variant >= SV_SELL1 && variant <= SV_SELL3) || (variant >= SV_SELLQ1 && variant <= SV_SELLQ3)) {
        switch (nvec) {
#define PCI_NV(n) case n: launch_sell<VT, XT, AT, n>(M, variant, val, X, ldx, Y, ldy, shift, s); break;
            PCI_NV(1) PCI_NV(2) PCI_NV(3) PCI_NV(4) PCI_NV(5) PCI_NV(6) PCI_NV(7) PCI_NV(8)
            PCI_NV(9) PCI_NV(10) PCI_NV(11) PCI_NV(12) PCI_NV(13) PCI_NV(14) PCI_NV(15) PCI_NV(16)
            PCI_NV(17) PCI_NV(18) PCI_NV(19) PCI_NV(20) PCI_NV(21) PCI_NV(22) PCI_NV(23) PCI_NV(24)
            PCI_NV(28) PCI_NV(32)
#undef PCI_NV
            default: throw Error(PCI_E_UNSUPPORTED, "SpMM block width not supported");
        }
        PCI_HIP(hipGetLastError());
        return;
    }
    switch (nvec) {
#define PCI_NV(n) case n: launch_nv<VT, XT, AT, n>(variant, grid, s, nd, rp, col, val, X, ldx, Y, ldy, shift, pp, part); break;
        PCI_NV(1) PCI_NV(2) PCI_NV(3) PCI_NV(4) PCI_NV(5) PCI_NV(6) PCI_NV(7) PCI_NV(8)
        PCI_NV(9) PCI_NV(10) PCI_NV(11) PCI_NV(12) PCI_NV(13) PCI_NV(14) PCI_NV(15) PCI_NV(16)
        PCI_NV(17) PCI_NV(18) PCI_NV(19) PCI_NV(20) PCI_NV(21) PCI_NV(22) PCI_NV(23) PCI_NV(24)
        PCI_NV(28) PCI_NV(32)
#undef PCI_NV
#define PCI_NV(n) case n: k_spmm<VT, XT, AT, n><<<grid, 256, 0, s>>>(nd, rp, col, val, X, ldx, Y, ldy, shift); break;
        PCI_NV(25) PCI_NV(26) PCI_NV(27) PCI_NV(29) PCI_NV(30) PCI_NV(31)
        PCI_NV(40) PCI_NV(48) PCI_NV(64)
#undef PCI_NV
        default: throw Error(PCI_E_UNSUPPORTED, "SpMM block width not supported");
    }
    PCI_HIP(hipGetLastError());
}

// The tuner builds every sliced copy it tries; keep only those a chosen
// kernel of this precision uses (a later tuning call rebuilds on demand).
static void release_unused_copies(Matrix &M, int prec, hipStream_t s) {
    if (M.blocks_only) return;  // the blocks are the matrix
    bool lds = false, wide = false, rb = false;
    for (int n = 0; n < 65; n++)
        for (int v = 0; v < 2; v++) {
            const int ch = M.spmm_choice[prec][n][v];
            rb |= ch >= SV_RBLK1 && ch <= SV_RB_LAST;
            lds |= (ch >= SV_SELL1 && ch <= SV_SELL3) || (ch >= SV_SELLQ1 && ch <= SV_SELLQ3);
            wide |= ch >= SV_SELLG1 && ch <= SV_SELLG2;
        }
    if ((!lds && M.sell[prec].built) || (!wide && M.sellg[prec].built)) PCI_HIP(hipStreamSynchronize(s));
    if (!lds && M.sell[prec].built) M.sell[prec].release();
    if (!wide && M.sellg[prec].built) M.sellg[prec].release();
    if (!rb && M.rblk[prec].built) {
        PCI_HIP(hipStreamSynchronize(s));
        M.rblk[prec].release();
    }
    if (!M.rblk[0].built && !M.rblk[1].built) {
        M.det_pos.release();
        M.slot_idx.release();
        M.slot.release();
    }
}

template <typename VT, typename XT, typename AT>
static void spmm_launch(Matrix &M, const VT *val, const XT *X, int ldx, XT *Y, int ldy, int nvec, AT shift,
                        bool padded, hipStream_t s) {
    constexpr int VW = sizeof(XT) == 4 ? 4 : 2;
    const bool aligned = ((uintptr_t)X % 16 == 0) && ((size_t)ldx * sizeof(XT)) % 16 == 0;
    const bool vec = vec_width(nvec) && aligned && (padded || nvec % VW == 0);
    signed char &choice = M.spmm_choice[sizeof(XT) == 4 ? 0 : 1][nvec][vec ? 1 : 0];
    // no CSR to fall back on: only the residual-block shapes, whose copies exist
    PCI_REQUIRE(!M.blocks_only || (vec && (nvec + VW - 1) / VW * VW <= 24), PCI_E_UNSUPPORTED,
                "H held as residual blocks only: block widths up to 24, 16-byte aligned rows");
    if (choice < 0) {
        PCI_REQUIRE(!M.noalloc || !vec, PCI_E_ARG,
                    "pci_gpu_spmm_noalloc: no SpMM kernel chosen yet for this block width and precision; call "
                    "pci_gpu_spmm (or a Davidson solve) with them first");
        if (!vec) {
            choice = SV_SCALAR;
        } else {
            // time each candidate on this call's operands (Y is rewritten below)
            hipEvent_t e0, e1;
            PCI_HIP(hipEventCreate(&e0));
            PCI_HIP(hipEventCreate(&e1));
            // per candidate: one untimed launch (builds its copy), one timed
            // launch, and two more only when that one is within 30 % of the
            // best so far (on large matrices the slow shapes take 0.1-0.5 s
            // a launch, so timing them once bounds the tuning cost)
            float best = 0.f;
            int bi = -1;
            auto timed = [&](int v, int reps) {
                PCI_HIP(hipEventRecord(e0, s));
                for (int r = 0; r < reps; r++) launch_variant<VT, XT, AT>(M, v, val, X, ldx, Y, ldy, nvec, shift, s);
                PCI_HIP(hipEventRecord(e1, s));
                PCI_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                PCI_HIP(hipEventElapsedTime(&ms, e0, e1));
                return ms / reps;
            };
            for (int v = SV_SCALAR; v < SV_COUNT; v++) {
                if (M.blocks_only && (v < SV_RBLK1 || v > SV_RB_LAST)) continue;
                launch_variant<VT, XT, AT>(M, v, val, X, ldx, Y, ldy, nvec, shift, s);
                float ms = timed(v, 1);
                if (bi < 0 || ms < 1.3f * best) ms = timed(v, 2);
                if (bi < 0 || ms < best) { best = ms; bi = v; }
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            choice = (signed char)bi;
            release_unused_copies(M, sizeof(XT) == 4 ? 0 : 1, s);
        }
    }
    launch_variant<VT, XT, AT>(M, choice, val, X, ldx, Y, ldy, nvec, shift, s);
}


// ---------------------------------------------------------------- head coupling
// Y[i, :] = sum over k < nd0 of H_ik X[k, :] for rows i >= nd0: the
// head-to-tail block of H applied to a block that lives on the first nd0
// rows.  Its elements are kept as a compact CSR of the tail rows (Matrix
// hc_*: columns ascending, the order of the full rows), made once per head
// size from the CSR's row prefixes or from the residual blocks; both give
// the same elements in the same order.  A tail row holds few of them (about
// 11 of 4,000 on C4, none on most rows), so one lane sums one row, reading
// the next element while the current one's X row arrives.
template <typename VT, typename XT, typename AT, int NV, int VW>
__global__ __launch_bounds__(256) void k_spmm_hc(int n, int row0, const int64_t *__restrict__ rp,
                                                 const int *__restrict__ col, const VT *__restrict__ val,
                                                 const XT *__restrict__ X, int ldx, XT *__restrict__ Y, int ldy) {
    constexpr int NVV = (NV + VW - 1) / VW;
    typedef XT V __attribute__((ext_vector_type(VW)));
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= n) return;
    AT acc[NVV * VW];
#pragma unroll
    for (int v = 0; v < NVV * VW; v++) acc[v] = 0;
    int64_t j = rp[t];
    const int64_t j1 = rp[t + 1];
    if (j < j1) {
        int k = col[j];
        AT a = (AT)val[j];
        while (true) {
            const bool more = j + 1 < j1;
            const int kn = more ? col[j + 1] : 0;
            const AT an = more ? (AT)val[j + 1] : (AT)0;
            const XT *xr = X + (size_t)k * ldx;
#pragma unroll
            for (int q = 0; q < NVV; q++) {
                const V x = *(const V *)(xr + q * VW);
#pragma unroll
                for (int v = 0; v < VW; v++) acc[q * VW + v] += a * (AT)x[v];
            }
            if (!more) break;
            k = kn;
            a = an;
            j++;
        }
    }
    const int row = row0 + t;
#pragma unroll
    for (int v = 0; v < NV; v++) Y[(size_t)row * ldy + v] = (XT)acc[v];
}

// cnt[t] = elements of tail row row0 + t with a column < ncol (CSR prefix)
__global__ void k_hc_count_csr(int n, int row0, int ncol, const int64_t *__restrict__ rp, const int *__restrict__ col,
                               int64_t *__restrict__ cnt) {
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= n) return;
    const int row = row0 + t;
    int64_t lo = rp[row], hi = rp[row + 1];
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (col[mid] < ncol) lo = mid + 1;
        else hi = mid;
    }
    cnt[t] = lo - rp[row];
}

__global__ void k_hc_fill_csr(int n, int row0, const int64_t *__restrict__ rp, const int *__restrict__ col,
                              const double *__restrict__ val, const int64_t *__restrict__ hrp,
                              int *__restrict__ hcol, double *__restrict__ h64, float *__restrict__ h32) {
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= n) return;
    const int64_t src = rp[row0 + t], dst = hrp[t], len = hrp[t + 1] - hrp[t];
    for (int64_t e = 0; e < len; e++) {
        const double v = val[src + e];
        hcol[dst + e] = col[src + e];
        h64[dst + e] = v;
        h32[dst + e] = (float)v;  // off-diagonal: the FP32 matrix's own rounding
    }
}

// number of a list's determinants below nd0 (the list is ascending)
__device__ inline int rb_head_count(const int *__restrict__ det, int n, int nd0) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (det[mid] < nd0) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// The same from the FP64 blocks: a tail determinant's elements with a head
// column are the nonzero values of its row in each of its lists over the
// list's head prefix.  FILL: written in ascending column order (insertion
// into the row's segment; the rows are short).
template <bool FILL>
__global__ void k_hc_blk(int n, int row0, int nd0, ResView R, const int *__restrict__ det_pos,
                         const int64_t *__restrict__ boff, const double *__restrict__ b64, int64_t *__restrict__ cnt,
                         const int64_t *__restrict__ hrp, int *__restrict__ hcol, double *__restrict__ h64,
                         float *__restrict__ h32) {
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= n) return;
    const int row = row0 + t;
    int64_t m = 0;
    const int64_t base = FILL ? hrp[t] : 0;
    for (int q = 0; q < R.nres; q++) {
        const int lid = R.det_res[(size_t)row * R.nres + q];
        const int p = det_pos[(size_t)row * R.nres + q];
        const int64_t lp = R.list_ptr[lid];
        const int ln = (int)(R.list_ptr[lid + 1] - lp);
        const int h = rb_head_count(R.list_det + lp, ln, nd0);
        const int steps = (ln + 7) >> 3;
        const double *blk = b64 + boff[lid];
        for (int c = 0; c < h; c++) {
            const double v = blk[rb_index(p, c, steps)];
            if (v == 0.0) continue;
            if (FILL) {
                const int k = R.list_det[lp + c];
                int64_t at = base + m;
                while (at > base && hcol[at - 1] > k) {
                    hcol[at] = hcol[at - 1];
                    h64[at] = h64[at - 1];
                    at--;
                }
                hcol[at] = k;
                h64[at] = v;
            }
            m++;
        }
    }
    if (!FILL) cnt[t] = m;
    else
        for (int64_t e = 0; e < m; e++) h32[base + e] = (float)h64[base + e];
}

void head_coupling_setup(Ctx &c, int nd0, hipStream_t s) {
    Matrix &M = c.H;
    if (M.hc_nd0 == nd0) return;
    M.hc_nd0 = -1;
    const int n = M.nd - nd0;
    if (n <= 0) {
        M.hc_nd0 = nd0;
        return;
    }
    const unsigned g = (unsigned)((n + 255) / 256);
    DBuf<int64_t> cnt;
    cnt.alloc(n);
    if (M.blocks_only) {
        k_hc_blk<false><<<g, 256, 0, s>>>(n, nd0, nd0, M.res, M.det_pos.p, M.rblk[1].boff.p,
                                          (const double *)M.rblk[1].val.p, cnt.p, nullptr, nullptr, nullptr, nullptr);
    } else {
        k_hc_count_csr<<<g, 256, 0, s>>>(n, nd0, nd0, M.rp.p, M.col.p, cnt.p);
    }
    PCI_HIP(hipGetLastError());
    M.hc_rp.alloc((size_t)n + 1);
    scan_exclusive(cnt.p, M.hc_rp.p, n, s, c.work);
    int64_t tot = 0;
    PCI_HIP(hipMemcpyAsync(&tot, M.hc_rp.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    PCI_HIP(hipStreamSynchronize(s));
    M.hc_col.alloc(tot);
    M.hc_v64.alloc(tot);
    M.hc_v32.alloc(tot);
    if (M.blocks_only) {
        k_hc_blk<true><<<g, 256, 0, s>>>(n, nd0, nd0, M.res, M.det_pos.p, M.rblk[1].boff.p,
                                         (const double *)M.rblk[1].val.p, nullptr, M.hc_rp.p, M.hc_col.p, M.hc_v64.p,
                                         M.hc_v32.p);
    } else {
        k_hc_fill_csr<<<g, 256, 0, s>>>(n, nd0, M.rp.p, M.col.p, M.val.p, M.hc_rp.p, M.hc_col.p, M.hc_v64.p,
                                        M.hc_v32.p);
    }
    PCI_HIP(hipGetLastError());
    PCI_HIP(hipStreamSynchronize(s));
    M.hc_nd0 = nd0;
}

template <typename VT, typename XT, typename AT>
static void hc_launch(const Matrix &M, const VT *val, int nd0, const XT *X, int ldx, XT *Y, int ldy, int nvec,
                      hipStream_t s) {
    constexpr int VW = sizeof(XT) == 4 ? 4 : 2;
    const int n = M.nd - nd0;
    if (n <= 0) return;
    PCI_REQUIRE(M.hc_nd0 == nd0, PCI_E_ARG, "head coupling not set up for this head size");
    const unsigned grid = (unsigned)((n + 255) / 256);
    switch (nvec) {
#define PCI_NV(k) case k: k_spmm_hc<VT, XT, AT, k, VW><<<grid, 256, 0, s>>>(n, nd0, M.hc_rp.p, M.hc_col.p, val, X, ldx, Y, ldy); break;
        PCI_NV(1) PCI_NV(2) PCI_NV(3) PCI_NV(4) PCI_NV(5) PCI_NV(6) PCI_NV(7) PCI_NV(8)
        PCI_NV(9) PCI_NV(10) PCI_NV(11) PCI_NV(12) PCI_NV(13) PCI_NV(14) PCI_NV(15) PCI_NV(16)
        PCI_NV(17) PCI_NV(18) PCI_NV(19) PCI_NV(20) PCI_NV(21) PCI_NV(22) PCI_NV(23) PCI_NV(24)
#undef PCI_NV
        default: throw Error(PCI_E_UNSUPPORTED, "head coupling: block width above 24");
    }
    PCI_HIP(hipGetLastError());
}

void head_couple_f64(const Ctx &c, int nd0, const double *X, int ldx, double *Y, int ldy, int nvec, hipStream_t s) {
    hc_launch<double, double, double>(c.H, c.H.hc_v64.p, nd0, X, ldx, Y, ldy, nvec, s);
}

void head_couple_f32(const Ctx &c, int nd0, const float *X, int ldx, float *Y, int ldy, int nvec, hipStream_t s) {
    hc_launch<float, float, float>(c.H, c.H.hc_v32.p, nd0, X, ldx, Y, ldy, nvec, s);
}

// --------------------------------------------------- Init4 block and diagonal
// Z(row, k) for the stored elements of rows < nd0 with k < nd0, diagonal
// shifted by -hmin (DiagInitApprox, src/pconf.F90:1876-1889); Z column-major.
__global__ __launch_bounds__(256) void k_dense_block(int nd0, const int64_t *__restrict__ rp,
                                                     const int *__restrict__ col, const double *__restrict__ val,
                                                     double hmin, double *__restrict__ Z) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= nd0) return;
    for (int64_t j = rp[row] + lane; j < rp[row + 1]; j += 64) {
        const int k = col[j];
        if (k < nd0) Z[(size_t)k * nd0 + row] = (k == row) ? val[j] - hmin : val[j];
    }
}

// D_i = H_ii - hmin (0 - hmin where the diagonal element is absent)
__global__ __launch_bounds__(256) void k_diag(int nd, const int64_t *__restrict__ rp, const int *__restrict__ col,
                                              const double *__restrict__ val, const int64_t *__restrict__ low,
                                              double hmin, double *__restrict__ D) {
    const int row = blockIdx.x * 256 + threadIdx.x;
    if (row >= nd) return;
    double d = 0.0;
    const int64_t l = low[row];
    if (l > 0) {
        const int64_t j = rp[row] + l - 1;
        if (col[j] == row) d = val[j];
    }
    D[row] = d - hmin;
}

// The same from the FP64 blocks: one wavefront per list, over the pairs of
// its head prefix (every stored pair of head determinants is in exactly
// one list, as a nonzero value; both orders are in the block).
__global__ __launch_bounds__(64) void k_dense_block_blk(int nd0, ResView R, const int64_t *__restrict__ boff,
                                                        const double *__restrict__ b64, double hmin,
                                                        double *__restrict__ Z) {
    const int lid = blockIdx.x;
    const int64_t lp = R.list_ptr[lid];
    const int n = (int)(R.list_ptr[lid + 1] - lp);
    const int h = rb_head_count(R.list_det + lp, n, nd0);
    const int steps = (n + 7) >> 3;
    const double *blk = b64 + boff[lid];
    for (int64_t e = threadIdx.x; e < (int64_t)h * h; e += 64) {
        const int p = (int)(e / h), q = (int)(e - (int64_t)p * h);
        const double v = blk[rb_index(p, q, steps)];
        if (v == 0.0) continue;
        const int i = R.list_det[lp + p], k = R.list_det[lp + q];
        Z[(size_t)k * nd0 + i] = (k == i) ? v - hmin : v;
    }
}

// A determinant's diagonal element is in its first list (the residual of its
// Ne-2 lowest positions) at its own position.
__global__ __launch_bounds__(256) void k_diag_blk(int nd, ResView R, const int *__restrict__ det_pos,
                                                  const int64_t *__restrict__ boff, const double *__restrict__ b64,
                                                  double hmin, double *__restrict__ D) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nd) return;
    const int lid = R.det_res[(size_t)i * R.nres], p = det_pos[(size_t)i * R.nres];
    const int n = (int)(R.list_ptr[lid + 1] - R.list_ptr[lid]);
    D[i] = b64[boff[lid] + rb_index(p, p, (n + 7) >> 3)] - hmin;
}

void head_block_and_diag(const Ctx &c, int nd0, double *Z, double *D, hipStream_t s) {
    const Matrix &M = c.H;
    if (M.blocks_only) {
        const RblkCopy &B = M.rblk[1];
        k_dense_block_blk<<<(unsigned)M.res.nlists, 64, 0, s>>>(nd0, M.res, B.boff.p, (const double *)B.val.p,
                                                                 M.hmin, Z);
        k_diag_blk<<<(unsigned)((M.nd + 255) / 256), 256, 0, s>>>(M.nd, M.res, M.det_pos.p, B.boff.p,
                                                                   (const double *)B.val.p, M.hmin, D);
    } else {
        k_dense_block<<<(unsigned)((nd0 + 3) / 4), 256, 0, s>>>(nd0, M.rp.p, M.col.p, M.val.p, M.hmin, Z);
        k_diag<<<(unsigned)((M.nd + 255) / 256), 256, 0, s>>>(M.nd, M.rp.p, M.col.p, M.val.p, M.lowlen.p, M.hmin,
                                                               D);
    }
    PCI_HIP(hipGetLastError());
}

void spmm_f64(Ctx &c, const double *X, int ldx, double *Y, int ldy, int nvec, hipStream_t s, bool padded) {
    Matrix &M = c.H;
    spmm_launch<double, double, double>(M, M.val.p, X, ldx, Y, ldy, nvec, -M.hmin, padded, s);
}

void spmm_f32(Ctx &c, const float *X, int ldx, float *Y, int ldy, int nvec, hipStream_t s, bool padded) {
    Matrix &M = c.H;
    // val32 carries the -hmin shift on its diagonal unless some diagonal
    // element is absent, in which case the shift is applied per row here
    spmm_launch<float, float, float>(M, M.val32.p, X, ldx, Y, ldy, nvec, M.shift32, padded, s);
}

void spmm_matrix_f64(Matrix &M, const double *X, int ldx, double *Y, int ldy, int nvec, double shift, hipStream_t s,
                     bool padded) {
    spmm_launch<double, double, double>(M, M.val.p, X, ldx, Y, ldy, nvec, shift, padded, s);
}

// The kernel that runs for (precision, nvec): a shape whose copy was found
// ineligible runs the CSR kernel k_spmm_xlds instead, and is reported as
// that (before its first use a shape is reported by its own name).
const char *spmm_variant_name(const Ctx &c, int prec, int nvec) {
    if (nvec < 1 || nvec > 64) return "";
    const int p = prec == 32 ? 0 : 1;
    const Matrix &M = c.H;
    for (int v = 1; v >= 0; v--) {
        const int ch = M.spmm_choice[p][nvec][v];
        if (ch < 0) continue;
        const bool sell = (ch >= SV_SELL1 && ch <= SV_SELL3) || (ch >= SV_SELLQ1 && ch <= SV_SELLQ3);
        if ((sell && !M.sell[p].eligible) || (ch >= SV_SELLG1 && ch <= SV_SELLG2 && !M.sellg[p].eligible) ||
            (ch >= SV_RBLK1 && ch <= SV_RB_LAST && !M.rblk[p].eligible))
            return kVariantName[SV_XLDS2];
        return kVariantName[ch];
    }
    return "";
}

// Algorithmic bytes of one application Y = (H - Hmin) X with the kernel in
// use for (nvec, precision): for the CSR and sliced kernels the CSR stream
// (value + 4-byte column per stored element, row pointers) plus X read and
// Y written once; for the residual-block kernel its own stream, which holds
// no column indices: every block value (zeros of the blocks included), per
// list entry its determinant index, the gathered X row and the slot written
// and read back, per determinant its list ids and positions, X (shift term)
// read and Y written once.
double spmm_bytes(const Ctx &c, int nvec, int prec) {
    const Matrix &M = c.H;
    const int p = prec == 32 ? 0 : 1;
    const double sv = prec == 32 ? 4.0 : 8.0;
    const int ch = nvec >= 1 && nvec <= 64 ? M.spmm_choice[p][nvec][1] : -1;
    if (ch >= SV_RBLK1 && ch <= SV_RB_LAST && M.rblk[p].built) {
        const double entries = (double)M.nd * M.res.nres;
        return (double)M.rblk[p].nvals * sv + entries * (4.0 + nvec * sv + 2.0 * nvec * sv + 8.0) +
               2.0 * M.nd * nvec * sv;
    }
    return (double)M.nnz_full * (sv + 4.0) + 8.0 * (M.nd + 1) + 2.0 * M.nd * nvec * sv;
}

// ------------------------------------------------ allocation-free SpMM
// Scratch of pci_gpu_spmm_noalloc: the partial sums of one product (the
// residual-block kernel's list slots, the sliced kernels' panel-group sums)
// in a file-scope device array that every GPU the library runs on holds
// from the library's load until the process ends.  Host code never uses
// the array's name as an address: scratch_setup looks up each device's
// copy once (hipGetSymbolAddress, that device current) and the kernels get
// that pointer.  Calls on one GPU share the array and leave the last call's
// contents; nothing in it needs to start at zero (every slot or partial sum
// a product reads is written by that product), and calls that overlap on
// one GPU race.
constexpr size_t NOALLOC_SCRATCH = (size_t)256 << 20;
static __device__ __align__(256) char scratch[NOALLOC_SCRATCH];
constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static char *g_scratch_ptr[MAX_DEVICES];

void scratch_setup(int device) {
    PCI_REQUIRE(device >= 0 && device < MAX_DEVICES, PCI_E_UNSUPPORTED, "device ordinal above 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void *p = nullptr;
    PCI_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    PCI_REQUIRE(p != nullptr, PCI_E_HIP, "hipGetSymbolAddress(scratch) returned no address");
    g_scratch_ptr[device] = (char *)p;
}

static char *scratch_ptr(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return device >= 0 && device < MAX_DEVICES ? g_scratch_ptr[device] : nullptr;
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
        PCI_REQUIRE((nvec >= 1 && nvec <= 32) || nvec == 40 || nvec == 48 || nvec == 64, PCI_E_ARG,
                    "nvec must be 1..32, 40, 48 or 64");
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
            spmm_f64(c, dx.p, ld, dy.p, ld, nvec, s, true);  // untimed: autotunes on first use
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
            spmm_f32(c, dx.p, ld, dy.p, ld, nvec, s, true);  // untimed: autotunes on first use
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

extern "C" int pci_gpu_spmm_variant(pci_ctx *ctx, int32_t nvec, int32_t precision, char *name, int32_t len) {
    try {
        PCI_REQUIRE(ctx && name && len > 0, PCI_E_ARG, "null argument");
        const char *v = spmm_variant_name(ctx->c, precision, nvec);
        snprintf(name, (size_t)len, "%s", v);
        return PCI_OK;
    } catch (const pci::Error &e) {
        set_last_error(e.what());
        return e.code;
    }
}

// Pin the kernel shape for (precision, nvec) on the resident matrix, or
// return the number of shapes (variant < 0).  Vector shapes apply to calls
// whose X rows are 16-byte aligned and padded; other calls keep k_spmm.
extern "C" int pci_gpu_spmm_set_variant(pci_ctx *ctx, int32_t nvec, int32_t precision, int32_t variant,
                                        int32_t *count) {
    try {
        PCI_REQUIRE(ctx, PCI_E_ARG, "null context");
        if (count) *count = SV_COUNT;
        if (variant < 0) return PCI_OK;
        PCI_REQUIRE(variant < SV_COUNT, PCI_E_ARG, "no such SpMM kernel shape");
        PCI_REQUIRE(nvec >= 1 && nvec <= 64 && (precision == 32 || precision == 64), PCI_E_ARG, "bad nvec/precision");
        PCI_REQUIRE(ctx->c.H.valid, PCI_E_ARG, "no Hamiltonian on the device");
        Matrix &M = ctx->c.H;
        PCI_REQUIRE(!M.blocks_only || (variant >= SV_RBLK1 && variant <= SV_RB_LAST), PCI_E_ARG,
                    "H is held as residual blocks only: just the residual-block shapes run");
        const int p = precision == 32 ? 0 : 1;
        M.spmm_choice[p][nvec][1] = (signed char)(vec_width(nvec) ? variant : SV_SCALAR);
        M.spmm_choice[p][nvec][0] = (signed char)SV_SCALAR;
        return PCI_OK;
    } catch (const pci::Error &e) {
        set_last_error(e.what());
        return e.code;
    }
}

extern "C" int pci_gpu_spmm_noalloc(pci_ctx *ctx, int32_t nvec, int32_t precision, const void *x, int32_t ldx,
                                    void *y, int32_t ldy) {
    try {
        PCI_REQUIRE(ctx && x && y, PCI_E_ARG, "null argument");
        Ctx &c = ctx->c;
        PCI_HIP(hipSetDevice(c.device));
        Matrix &M = c.H;
        PCI_REQUIRE(M.valid, PCI_E_ARG, "no Hamiltonian on the device");
        PCI_REQUIRE(nvec >= 1 && nvec <= 64, PCI_E_ARG, "nvec out of range");
        PCI_REQUIRE(precision == 32 || precision == 64, PCI_E_ARG, "precision must be 32 or 64");
        const int vw = precision == 32 ? 4 : 2, es = precision == 32 ? 4 : 8;
        const int need_ld = (nvec + vw - 1) / vw * vw;
        PCI_REQUIRE(ldx >= need_ld && ldy >= need_ld, PCI_E_ARG,
                    "pci_gpu_spmm_noalloc: ldx and ldy must be at least nvec rounded up to " + std::to_string(vw));
        PCI_REQUIRE((uintptr_t)x % 16 == 0 && (uintptr_t)y % 16 == 0 && ((size_t)ldx * es) % 16 == 0 &&
                        ((size_t)ldy * es) % 16 == 0,
                    PCI_E_ARG, "pci_gpu_spmm_noalloc: x and y rows must be 16-byte aligned");
        PCI_REQUIRE(precision == 64 || M.has32, PCI_E_ARG,
                    "pci_gpu_spmm_noalloc: the FP32 values are not built; call pci_gpu_spmm with precision 32 first");
        char *sp = scratch_ptr(c.device);
        PCI_REQUIRE(sp != nullptr, PCI_E_ARG, "scratch of this device not set up (context not created on it)");
        struct Guard {
            Matrix &M;
            ~Guard() {
                M.noalloc = false;
                M.ext = nullptr;
                M.ext_cap = 0;
            }
        } guard{M};
        M.noalloc = true;
        M.ext = sp;
        M.ext_cap = NOALLOC_SCRATCH;
        hipStream_t s = c.stream;
        if (precision == 64) spmm_f64(c, (const double *)x, ldx, (double *)y, ldy, nvec, s, true);
        else spmm_f32(c, (const float *)x, ldx, (float *)y, ldy, nvec, s, true);
        PCI_HIP(hipGetLastError());
        PCI_HIP(hipStreamSynchronize(s));
        return PCI_OK;
    } catch (const pci::Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return PCI_E_HIP;
    }
}
