// CSF basis (kCSF = 1) on the device: the basis, H in it, and the way back.
//
// Reference: jbasis (src/csf.f90:17-407) with build_j2 (409-452),
// formh_sym with compute_zzc_block and Hmatrix (538-1051), idif
// (1189-1227), unsym (1088-1134); pconf runs them at src/pconf.F90:146-149,
// 211-214 and 247-255.
//
// Basis.  The configurations of one class (nonequiv_conf) have the same
// determinant pattern, so one J^2 block per class is diagonalised, built from
// the class's first configuration.  Blocks of equal size are built by one
// kernel and solved by one batched rocSOLVER dsyevd; the eigenvalues come back
// to the host, which checks each J and picks the J = Jm columns (ascending
// eigenvalue order, as jbasis keeps them); a gather kernel packs them into
// the resident coefficients ccj, class after class.
//
// H.  H_csf = C^T H C is block-sparse: configurations I >= J that differ by at
// most two electrons give a dense nc(I) x nc(J) block.  Pairs are found on
// the host from the configurations' shell lists: two configurations differ by
// at most two electrons exactly when they share an (Ne-2)-electron residual,
// so each configuration is filed under its residuals and pairs are read off
// the residual lists (rather than testing all Nc^2 / 2 pairs as idif does).
// A one-wavefront workgroup computes four blocks in turn (in chunks of at
// most 16 x 16 CSFs): it walks the determinants n of I in tiles of 16 and
// those of J in chunks of 64, screens the 16 x 64 pairs for at most two
// differences (Rspq's; the others give zero), evaluates the Slater-Condon
// elements <n|H|k> (Hmatrix) of the listed candidates only -- packed, so the
// lanes all do element work -- into LDS (about 16 KB per workgroup, so ten
// or more independent workgroups share a CU),
// and accumulates B(n, jc) = sum_k h(n, k) C_J(k, jc) in k order and then
// Z(ic, jc) += B(n, jc) C_I(n, ic) in n order -- compute_zzc_block's order
// term by term, with contraction off, so every value equals the CPU
// oracle's on the same basis bit for bit.  The block is written into both
// triangles of the resident full-storage CSR, whose layout is known before
// any value is: every row of configuration I holds, for each neighbour J
// ascending, nc(J) columns.
//
// unsym: y(n) = sum over the CSFs ic of n's configuration, in order, of
// x(mdcs + ic) C(n, ic), one thread per (vector, determinant).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <string>
#include <thread>
#include <unordered_map>

#include "internal.h"

namespace pci {

namespace {

constexpr int CSF_WG = 64;    // one wavefront per workgroup (its barriers cost nothing)
constexpr int CSF_NT = 16;    // determinants n per tile
constexpr int CSF_KT = 64;    // determinants k per chunk
constexpr int CSF_JC = 16;    // CSF columns jc per chunk
constexpr int CSF_IC = 16;    // CSF rows ic per chunk
constexpr int CSF_TPW = 4;    // configuration pairs per workgroup (in order)
constexpr int CSF_BQ = CSF_NT * CSF_JC / CSF_WG;  // B entries per lane

struct CsfConf {
    int det0, ndc, nc, mdcs;
    int64_t place, rp0, rowlen;
};
struct CsfPair {
    int i, j, offij, offji;
};

// J^2 blocks of a group of classes with m determinants each: block g is
// column-major m x m at A + g m^2, over the determinants det0[g] ..
// det0[g] + m - 1.  Entry (r, s) is build_j2's value of the pair (max(r,s),
// min(r,s)) -- F_J2 of Rspq's differences, zero for more than two -- and
// stays +0 where build_j2 leaves zz untouched.
template <int W>
__global__ __launch_bounds__(256) void k_csf_j2(Tables T, const uint64_t *__restrict__ occ, int nd, int m,
                                                const int *__restrict__ det0, double *__restrict__ A) {
    const int g = blockIdx.y;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)m * m) return;
    const int r = (int)(e % m), s = (int)(e / m);
    const int n = r > s ? r : s, k = r > s ? s : r;
    const int dn = det0[g] + n, dk = det0[g] + k;
    uint64_t a[W], b[W];
    int pc = 0;
#pragma unroll
    for (int w = 0; w < W; w++) {
        a[w] = occ[(size_t)w * nd + dn];
        b[w] = occ[(size_t)w * nd + dk];
        pc += popc64(a[w] ^ b[w]);
    }
    double t = 0.0;
    if (pc == 0 || pc == 4) t = j2_element<W>(T, a, b, pc >> 1);
    A[(int64_t)g * m * m + e] = (t != 0.0) ? t : 0.0;
}

// ccj[dst[i] .. dst[i] + m) = src[so[i] .. so[i] + m): the selected columns
__global__ void k_csf_gather(const double *__restrict__ src, int m, const int64_t *__restrict__ so,
                             const int64_t *__restrict__ dst, double *__restrict__ ccj) {
    const int i = blockIdx.x;
    for (int r = threadIdx.x; r < m; r += blockDim.x) ccj[dst[i] + r] = src[so[i] + r];
}

template <int W>
__global__ __launch_bounds__(CSF_WG) void k_csf_h(Tables T, const uint64_t *__restrict__ occ, int nd, int npair,
                                               const CsfPair *__restrict__ pairs, const CsfConf *__restrict__ conf,
                                               const double *__restrict__ ccj, int *__restrict__ col,
                                               double *__restrict__ val) {
#pragma clang fp contract(off)
    __shared__ double hs[CSF_NT * CSF_KT];
    __shared__ double bs[CSF_NT * CSF_JC];
    __shared__ double zs[CSF_IC * CSF_JC];
    __shared__ short cand[CSF_NT * CSF_KT];
    __shared__ int ncand;
    __shared__ uint64_t oa[CSF_NT * W], ob[CSF_KT * W];  // the tile's determinants, word-major
    const int tid = threadIdx.x;
    const int t1 = min(npair, (blockIdx.x + 1) * CSF_TPW);
    for (int t = blockIdx.x * CSF_TPW; t < t1; t++) {
        const CsfPair pr = pairs[t];
        const CsfConf ci = conf[pr.i], cj = conf[pr.j];
        const int ndi = ci.ndc, ndj = cj.ndc;
        const int njch = (cj.nc + CSF_JC - 1) / CSF_JC, nch = ((ci.nc + CSF_IC - 1) / CSF_IC) * njch;
        for (int ch = 0; ch < nch; ch++) {
            const int ic0 = (ch / njch) * CSF_IC, jc0 = (ch % njch) * CSF_JC;
            // on the diagonal block only jc <= ic is stored
            if (pr.i == pr.j && jc0 > min(ci.nc, ic0 + CSF_IC) - 1) continue;
            const int icn = min(CSF_IC, ci.nc - ic0), jcn = min(CSF_JC, cj.nc - jc0);
            const double *__restrict__ Ci = ccj + ci.place + (int64_t)ic0 * ndi;
            const double *__restrict__ Cj = ccj + cj.place + (int64_t)jc0 * ndj;
            for (int e = tid; e < icn * jcn; e += CSF_WG) zs[e] = 0.0;
            for (int n0 = 0; n0 < ndi; n0 += CSF_NT) {
                const int nt = min(CSF_NT, ndi - n0);
                // this lane's B entries (n0 + e / jcn, jc0 + e % jcn), e = tid + CSF_WG q
                double b[CSF_BQ];
#pragma unroll
                for (int q = 0; q < CSF_BQ; q++) b[q] = 0.0;
                for (int k0 = 0; k0 < ndj; k0 += CSF_KT) {
                    const int kt = min(CSF_KT, ndj - k0);
                    // screen the tile's pairs (at most two differences) and list
                    // the candidates, so that the element evaluations run packed
                    if (tid == 0) ncand = 0;
                    for (int e = tid; e < (nt + kt) * W; e += CSF_WG) {
                        const int w = e % W, r = e / W;
                        if (r < nt) oa[w * CSF_NT + r] = occ[(size_t)w * nd + ci.det0 + n0 + r];
                        else ob[w * CSF_KT + r - nt] = occ[(size_t)w * nd + cj.det0 + k0 + r - nt];
                    }
                    __syncthreads();
                    for (int e = tid; e < nt * kt; e += CSF_WG) {
                        const int nl = e / kt, kl = e - nl * kt;
                        int pc = 0;
#pragma unroll
                        for (int w = 0; w < W; w++) pc += popc64(oa[w * CSF_NT + nl] ^ ob[w * CSF_KT + kl]);
                        if (pc <= 4) cand[atomicAdd(&ncand, 1)] = (short)(nl * CSF_KT + kl);
                        else hs[nl * CSF_KT + kl] = 0.0;
                    }
                    __syncthreads();
                    for (int i = tid; i < ncand; i += CSF_WG) {
                        const int slot = cand[i];
                        const int nl = slot / CSF_KT, kl = slot - nl * CSF_KT;
                        const int dk = cj.det0 + k0 + kl;
                        uint64_t A[W], B[W];
                        int pc = 0;
#pragma unroll
                        for (int w = 0; w < W; w++) {
                            A[w] = oa[w * CSF_NT + nl];
                            B[w] = ob[w * CSF_KT + kl];
                            pc += popc64(A[w] ^ B[w]);
                        }
                        // Hmatrix(idet1 = n, idet2 = k) with E_k = Diag(k) (src/csf.f90:988-989)
                        const double ek = T.kdsig ? T.dg[dk] : 0.0;
                        hs[slot] = element<W>(T, A, B, pc >> 1, ek);
                    }
                    __syncthreads();
#pragma unroll
                    for (int q = 0; q < CSF_BQ; q++) {
                        const int e = tid + CSF_WG * q;
                        if (e < nt * jcn) {
                            const int nl = e / jcn, jl = e - nl * jcn;
                            const double *__restrict__ cc = Cj + (int64_t)jl * ndj + k0;
                            double acc = b[q];
                            for (int kl = 0; kl < kt; kl++) {
                                const double h = hs[nl * CSF_KT + kl];
                                if (fabs(h) < 1e-20) continue;
                                acc = acc + h * cc[kl];
                            }
                            b[q] = acc;
                        }
                    }
                    __syncthreads();
                }
#pragma unroll
                for (int q = 0; q < CSF_BQ; q++) {
                    const int e = tid + CSF_WG * q;
                    if (e < nt * jcn) {
                        const int nl = e / jcn, jl = e - nl * jcn;
                        bs[nl * CSF_JC + jl] = b[q];
                    }
                }
                __syncthreads();
                for (int e = tid; e < icn * jcn; e += CSF_WG) {
                    const int il = e / jcn, jl = e - il * jcn;
                    const double *__restrict__ cc = Ci + (int64_t)il * ndi + n0;
                    double z = zs[e];
                    for (int nl = 0; nl < nt; nl++) z = z + bs[nl * CSF_JC + jl] * cc[nl];
                    zs[e] = z;
                }
                __syncthreads();
            }
            for (int e = tid; e < icn * jcn; e += CSF_WG) {
                const int il = e / jcn, jl = e - il * jcn;
                const int ic = ic0 + il, jc = jc0 + jl;
                const double z = zs[e];
                if (pr.i != pr.j) {
                    int64_t p = ci.rp0 + (int64_t)ic * ci.rowlen + pr.offij + jc;
                    col[p] = cj.mdcs + jc;
                    val[p] = z;
                    p = cj.rp0 + (int64_t)jc * cj.rowlen + pr.offji + ic;
                    col[p] = ci.mdcs + ic;
                    val[p] = z;
                } else if (jc <= ic) {
                    int64_t p = ci.rp0 + (int64_t)ic * ci.rowlen + pr.offij + jc;
                    col[p] = ci.mdcs + jc;
                    val[p] = z;
                    if (jc != ic) {
                        p = ci.rp0 + (int64_t)jc * ci.rowlen + pr.offij + ic;
                        col[p] = ci.mdcs + ic;
                        val[p] = z;
                    }
                }
            }
            __syncthreads();  // zs is cleared for the next chunk
        }
    }
}

// numzero and Hmin over the stored lower triangle: acc[0] += zeros, bmin[block]
__global__ __launch_bounds__(256) void k_csf_stats(int nrow, const int64_t *__restrict__ rp,
                                                   const int64_t *__restrict__ low, const double *__restrict__ val,
                                                   unsigned long long *__restrict__ acc, double *__restrict__ bmin) {
    __shared__ unsigned long long s_z[256];
    __shared__ double s_m[256];
    unsigned long long z = 0;
    double m = DBL_MAX;
    for (int r = blockIdx.x * 256 + threadIdx.x; r < nrow; r += gridDim.x * 256)
        for (int64_t p = rp[r]; p < rp[r] + low[r]; p++) {
            const double v = val[p];
            if (v == 0.0) z++;
            m = v < m ? v : m;
        }
    s_z[threadIdx.x] = z;
    s_m[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            s_z[threadIdx.x] += s_z[threadIdx.x + o];
            s_m[threadIdx.x] = s_m[threadIdx.x + o] < s_m[threadIdx.x] ? s_m[threadIdx.x + o] : s_m[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(&acc[0], s_z[0]);
        bmin[blockIdx.x] = s_m[0];
    }
}

__global__ __launch_bounds__(256) void k_csf_unsym(int nd, int nvec, int ncsf, const int *__restrict__ det_conf,
                                                   const CsfConf *__restrict__ conf, const double *__restrict__ ccj,
                                                   const double *__restrict__ x, double *__restrict__ y) {
#pragma clang fp contract(off)
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)nd * nvec) return;
    const int v = (int)(e / nd), n = (int)(e - (int64_t)v * nd);
    const CsfConf cf = conf[det_conf[n]];
    const int id = n - cf.det0;
    double acc = 0.0;
    for (int ic = 0; ic < cf.nc; ic++)
        acc = acc + x[(int64_t)v * ncsf + cf.mdcs + ic] * ccj[cf.place + (int64_t)ic * cf.ndc + id];
    y[e] = acc;
}

template <int W>
void launch_j2(const Ctx &c, const Tables &T, int m, int g, const int *det0, double *A) {
    dim3 grid((unsigned)(((int64_t)m * m + 255) / 256), (unsigned)g);
    k_csf_j2<W><<<grid, 256, 0, c.stream>>>(T, c.occ.p, c.nd, m, det0, A);
}

template <int W>
void launch_h(const Ctx &c, const Tables &T, int npair, const CsfPair *pairs, const CsfConf *conf,
              const double *ccj, int *col, double *val) {
    const unsigned grid = (unsigned)((npair + CSF_TPW - 1) / CSF_TPW);
    k_csf_h<W><<<grid, CSF_WG, 0, c.stream>>>(T, c.occ.p, c.nd, npair, pairs, conf, ccj, col, val);
}

#define PCI_CSF_W(F, ...)                                                                          \
    switch (c.W) {                                                                                 \
        case 1: F<1>(__VA_ARGS__); break;   case 2: F<2>(__VA_ARGS__); break;                     \
        case 3: F<3>(__VA_ARGS__); break;   case 4: F<4>(__VA_ARGS__); break;                     \
        case 5: F<5>(__VA_ARGS__); break;   case 6: F<6>(__VA_ARGS__); break;                     \
        case 7: F<7>(__VA_ARGS__); break;   case 8: F<8>(__VA_ARGS__); break;                     \
        case 9: F<9>(__VA_ARGS__); break;   case 10: F<10>(__VA_ARGS__); break;                   \
        case 11: F<11>(__VA_ARGS__); break; case 12: F<12>(__VA_ARGS__); break;                   \
        case 13: F<13>(__VA_ARGS__); break; case 14: F<14>(__VA_ARGS__); break;                   \
        case 15: F<15>(__VA_ARGS__); break; case 16: F<16>(__VA_ARGS__); break;                   \
        default: throw Error(PCI_E_UNSUPPORTED, "unsupported word count");                         \
    }

rocblas_handle blas_of(Ctx &c) {
    if (!c.blas) {
        rocblas_handle h = nullptr;
        PCI_REQUIRE(rocblas_create_handle(&h) == rocblas_status_success, PCI_E_HIP, "rocblas_create_handle failed");
        c.blas = h;
    }
    rocblas_set_stream((rocblas_handle)c.blas, c.stream);
    return (rocblas_handle)c.blas;
}

// per-configuration records of the basis (det0, ndc, nc, mdcs, place)
std::vector<CsfConf> conf_records(const Ctx &c) {
    const auto &B = c.csf;
    std::vector<CsfConf> cf(c.nc);
    for (int i = 0; i < c.nc; i++) {
        const int q = B.cls[i];
        cf[i].det0 = c.h_conf_start[i];
        cf[i].ndc = c.h_conf_start[i + 1] - c.h_conf_start[i];
        cf[i].nc = B.ndcs[q];
        cf[i].mdcs = B.mdcs[i];
        cf[i].place = B.place[q];
        cf[i].rp0 = 0;
        cf[i].rowlen = 0;
    }
    return cf;
}

// Lower neighbours of each configuration with CSFs: J <= I, J with CSFs,
// differing from I by at most two electrons (idif <= 2), ascending.
std::vector<std::vector<int>> lower_neighbours(const Ctx &c, const std::vector<int> &nc) {
    const int ncf = c.nc, ne = c.ne;
    std::vector<std::vector<int>> nb(ncf);
    if (ne <= 2) {
        for (int i = 0; i < ncf; i++)
            if (nc[i] > 0)
                for (int j = 0; j <= i; j++)
                    if (nc[j] > 0) nb[i].push_back(j);
        return nb;
    }
    // residual (shells left after taking two electrons away) -> configurations
    std::unordered_map<std::string, std::vector<int>> lists;
    std::vector<std::vector<std::string>> keys(ncf);
    std::string key((size_t)(ne - 2) * sizeof(int), '\0');
    for (int i = 0; i < ncf; i++) {
        if (nc[i] == 0) continue;
        const int *sh = c.h_conf_shells.data() + (size_t)i * ne;
        for (int a = 0; a < ne; a++)
            for (int b = a + 1; b < ne; b++) {
                // equal shells give equal residuals: take the first of a run only
                if (a > 0 && sh[a] == sh[a - 1]) continue;
                if (b > a + 1 && sh[b] == sh[b - 1]) continue;
                int q = 0;
                for (int t = 0; t < ne; t++)
                    if (t != a && t != b) memcpy(&key[(size_t)q++ * sizeof(int)], &sh[t], sizeof(int));
                auto &l = lists[key];
                if (l.empty() || l.back() != i) l.push_back(i);
                keys[i].push_back(key);
            }
    }
    // the residual lists of each configuration, then the lower neighbours,
    // configurations dealt round-robin to host threads
    std::vector<std::vector<const std::vector<int> *>> lref(ncf);
    for (int i = 0; i < ncf; i++)
        for (const auto &k : keys[i]) lref[i].push_back(&lists.find(k)->second);
    const int nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (int t = 0; t < nth; t++)
        th.emplace_back([&, t] {
            std::vector<int> mark(ncf, -1);
            for (int i = t; i < ncf; i += nth) {
                for (const auto *l : lref[i])
                    for (int j : *l) {
                        if (j > i) break;  // lists are ascending
                        if (mark[j] != i) {
                            mark[j] = i;
                            nb[i].push_back(j);
                        }
                    }
                std::sort(nb[i].begin(), nb[i].end());
            }
        });
    for (auto &x : th) x.join();
    return nb;
}

// The classes of the configurations (0-based, B.cls) and the determinants of
// each class (B.ndc_neq), checked: every configuration of a class has as
// many determinants as the first.  Returns the first configuration of each
// class (-1 for none).  Leaves the basis invalid.
std::vector<int> set_classes(Ctx &c, int nconf_neq, const int32_t *nc_neq) {
    auto &B = c.csf;
    B.valid = false;
    const int nc = c.nc;
    std::vector<int> first(nconf_neq, -1);
    B.cls.assign(nc, 0);
    B.ndc_neq.assign(nconf_neq, 0);
    for (int i = 0; i < nc; i++) {
        PCI_REQUIRE(nc_neq[i] >= 1 && nc_neq[i] <= nconf_neq, PCI_E_ARG, "configuration class out of range");
        const int q = nc_neq[i] - 1, ndi = c.h_conf_start[i + 1] - c.h_conf_start[i];
        B.cls[i] = q;
        if (first[q] < 0) {
            first[q] = i;
            B.ndc_neq[q] = ndi;
        }
        PCI_REQUIRE(ndi == B.ndc_neq[q], PCI_E_ARG,
                    "configurations of one class have different numbers of determinants");
    }
    B.nneq = nconf_neq;
    return first;
}

// From B.ndcs: each class's first coefficient (iplace_cj, src/csf.f90:
// 306-310), each configuration's first CSF (mdcs, 372-381), Ncsf; returns
// nccj.
int64_t set_layout(Ctx &c) {
    auto &B = c.csf;
    B.place.assign(B.nneq, 0);
    int64_t tot = 0;
    for (int q = 0; q < B.nneq; q++) {
        B.place[q] = tot;
        tot += (int64_t)B.ndcs[q] * B.ndc_neq[q];
    }
    B.nccj = tot;
    B.mdcs.assign(c.nc, 0);
    int n = 0;
    for (int i = 0; i < c.nc; i++) {
        B.mdcs[i] = n;
        n += B.ndcs[B.cls[i]];
    }
    B.ncsf = n;
    return tot;
}

}  // namespace
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

extern "C" int pci_gpu_csf_basis(pci_ctx *ctx, int32_t nconf_neq, const int32_t *nc_neq, int32_t *ndcs,
                                 int32_t *ncsf, int64_t *nccj) {
    API_BEGIN
    PCI_REQUIRE(ctx, PCI_E_ARG, "null context");
    Ctx &c = ctx->c;
    PCI_HIP(hipSetDevice(c.device));
    PCI_REQUIRE(c.have_space, PCI_E_ARG, "pci_gpu_csf_basis: set the space first");
    PCI_REQUIRE(nconf_neq >= 1 && nc_neq, PCI_E_ARG, "pci_gpu_csf_basis: no configuration classes");
    auto &B = c.csf;
    const int nc = c.nc;
    std::vector<int> first = set_classes(c, nconf_neq, nc_neq);
    hipStream_t s = c.stream;
    Tables T = c.tables();
    rocblas_handle h = blas_of(c);
    // classes by block size, one batched solve per size
    std::map<int, std::vector<int>> groups;
    for (int q = 0; q < nconf_neq; q++)
        if (first[q] >= 0 && B.ndc_neq[q] > 0) groups[B.ndc_neq[q]].push_back(q);
    B.ndcs.assign(nconf_neq, 0);
    std::vector<std::vector<int>> selcol(nconf_neq);  // selected columns of each class's eigenvectors
    // all groups' blocks in one buffer (group after group), kept until the
    // selected columns are gathered
    int64_t atot = 0;
    for (auto &kv : groups) atot += (int64_t)kv.second.size() * kv.first * kv.first;
    DBuf<double> Aall;
    Aall.alloc((size_t)std::max<int64_t>(atot, 1));
    std::vector<int64_t> gbase;
    int64_t abase = 0;
    for (auto &kv : groups) {
        const int m = kv.first, g = (int)kv.second.size();
        gbase.push_back(abase);
        double *A = Aall.p + abase;
        abase += (int64_t)g * m * m;
        std::vector<int> det0(g);
        for (int i = 0; i < g; i++) det0[i] = c.h_conf_start[first[kv.second[i]]];
        DBuf<int> d0, info;
        DBuf<double> D, E;
        upload(d0, det0.data(), g, s);
        D.alloc((size_t)g * m);
        E.alloc((size_t)g * m);
        info.alloc(g);
        PCI_CSF_W(launch_j2, c, T, m, g, d0.p, A);
        PCI_HIP(hipGetLastError());
        const rocblas_status st =
            rocsolver_dsyevd_strided_batched(h, rocblas_evect_original, rocblas_fill_upper, m, A, m,
                                             (rocblas_stride)m * m, D.p, m, E.p, m, info.p, g);
        PCI_REQUIRE(st == rocblas_status_success, PCI_E_HIP, "rocsolver_dsyevd_strided_batched failed");
        std::vector<double> de((size_t)g * m);
        std::vector<int> inf(g);
        PCI_HIP(hipMemcpyAsync(de.data(), D.p, de.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        PCI_HIP(hipMemcpyAsync(inf.data(), info.p, g * sizeof(int), hipMemcpyDeviceToHost, s));
        PCI_HIP(hipStreamSynchronize(s));
        for (int i = 0; i < g; i++) {
            const int q = kv.second[i];
            PCI_REQUIRE(inf[i] == 0, PCI_E_HIP, "jbasis: the J^2 eigensolver failed");
            for (int r = 0; r < m; r++) {
                // tj = (sqrt(1 + de) - 1) / 2, jtt = int(2 tj + 1e-4) (src/csf.f90:175-183)
                const double tj = 0.5 * (std::sqrt(1.0 + de[(size_t)i * m + r]) - 1.0);
                const int jtt = (int)(2.0 * tj + 0.0001);
                PCI_REQUIRE(std::fabs(2.0 * tj - jtt) <= 1e-7, PCI_E_ARG,
                            "Value of j in jbasis is wrong: J=" + std::to_string(tj) + " in configuration " +
                                std::to_string(first[q] + 1));
                if (jtt == c.mj) selcol[q].push_back(r);
            }
            B.ndcs[q] = (int)selcol[q].size();
        }
    }
    const int64_t tot = set_layout(c);
    B.ccj.alloc((size_t)std::max<int64_t>(tot, 1));
    std::vector<int64_t> so, dst;
    size_t gi = 0;
    for (auto &kv : groups) {
        const int m = kv.first;
        for (size_t i = 0; i < kv.second.size(); i++) {
            const int q = kv.second[i];
            for (size_t k = 0; k < selcol[q].size(); k++) {
                so.push_back(gbase[gi] + ((int64_t)i * m + selcol[q][k]) * m);
                dst.push_back(B.place[q] + (int64_t)k * m);
            }
        }
        if (!so.empty()) {
            DBuf<int64_t> dso, ddst;
            upload(dso, so.data(), so.size(), s);
            upload(ddst, dst.data(), dst.size(), s);
            k_csf_gather<<<(unsigned)so.size(), 64, 0, s>>>(Aall.p, m, dso.p, ddst.p, B.ccj.p);
            PCI_HIP(hipGetLastError());
            PCI_HIP(hipStreamSynchronize(s));
        }
        so.clear();
        dst.clear();
        gi++;
    }
    B.valid = true;
    if (ndcs)
        for (int q = 0; q < nconf_neq; q++) ndcs[q] = B.ndcs[q];
    if (ncsf) *ncsf = B.ncsf;
    if (nccj) *nccj = tot;
    API_END
}

extern "C" int pci_gpu_csf_set_basis(pci_ctx *ctx, int32_t nconf_neq, const int32_t *nc_neq, const int32_t *ndcs,
                                     const double *ccj) {
    API_BEGIN
    PCI_REQUIRE(ctx, PCI_E_ARG, "null context");
    Ctx &c = ctx->c;
    PCI_HIP(hipSetDevice(c.device));
    PCI_REQUIRE(c.have_space, PCI_E_ARG, "pci_gpu_csf_set_basis: set the space first");
    PCI_REQUIRE(nconf_neq >= 1 && nc_neq && ndcs, PCI_E_ARG, "pci_gpu_csf_set_basis: no configuration classes");
    auto &B = c.csf;
    set_classes(c, nconf_neq, nc_neq);
    B.ndcs.assign(nconf_neq, 0);
    for (int q = 0; q < nconf_neq; q++) {
        PCI_REQUIRE(ndcs[q] >= 0 && ndcs[q] <= B.ndc_neq[q], PCI_E_ARG,
                    "pci_gpu_csf_set_basis: more CSFs than determinants in a class");
        B.ndcs[q] = ndcs[q];
    }
    const int64_t tot = set_layout(c);
    PCI_REQUIRE(tot == 0 || ccj, PCI_E_ARG, "pci_gpu_csf_set_basis: null coefficients");
    upload(B.ccj, ccj, (size_t)tot, c.stream);
    PCI_HIP(hipStreamSynchronize(c.stream));
    B.valid = true;
    API_END
}

extern "C" int pci_gpu_csf_get_basis(pci_ctx *ctx, double *ccj) {
    API_BEGIN
    PCI_REQUIRE(ctx && ccj, PCI_E_ARG, "null argument");
    Ctx &c = ctx->c;
    PCI_REQUIRE(c.csf.valid, PCI_E_ARG, "no CSF basis (pci_gpu_csf_basis)");
    PCI_HIP(hipSetDevice(c.device));
    if (c.csf.nccj > 0) {
        PCI_HIP(hipMemcpyAsync(ccj, c.csf.ccj.p, c.csf.nccj * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        PCI_HIP(hipStreamSynchronize(c.stream));
    }
    API_END
}

extern "C" int pci_gpu_build_h_csf(pci_ctx *ctx, int64_t *numh, int64_t *numzero, double *hmin) {
    API_BEGIN
    PCI_REQUIRE(ctx, PCI_E_ARG, "null context");
    Ctx &c = ctx->c;
    PCI_HIP(hipSetDevice(c.device));
    PCI_REQUIRE(c.csf.valid, PCI_E_ARG, "pci_gpu_build_h_csf: build the CSF basis first (pci_gpu_csf_basis)");
    PCI_REQUIRE(c.have_integrals, PCI_E_ARG, "pci_gpu_build_h_csf: set the integrals first");
    PCI_REQUIRE(c.ksig < 2 || c.have_sigma2, PCI_E_ARG, "Ksig = 2 needs pci_gpu_set_sigma2 first");
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t s = c.stream;
    const auto &B = c.csf;
    const int ncf = c.nc, ncsf = B.ncsf;
    std::vector<CsfConf> cf = conf_records(c);
    std::vector<int> nc(ncf);
    for (int i = 0; i < ncf; i++) nc[i] = cf[i].nc;
    auto low = lower_neighbours(c, nc);
    // row layout: every row of configuration i holds its lower neighbours'
    // blocks (ascending, i itself last), then its upper ones (ascending)
    std::vector<int64_t> lowsum(ncf, 0), upsum(ncf, 0), diagoff(ncf, 0);
    int64_t npair = 0;
    for (int i = 0; i < ncf; i++)
        for (int j : low[i]) {
            if (j == i) diagoff[i] = lowsum[i];
            lowsum[i] += nc[j];
            if (j != i) upsum[j] += nc[i];
            npair++;
        }
    std::vector<int64_t> rp(ncsf + 1, 0), lowlen(ncsf, 0);
    int64_t pos = 0, numh_ = 0, ncand = 0;
    for (int i = 0; i < ncf; i++) {
        if (nc[i] == 0) continue;
        const int64_t rl = lowsum[i] + upsum[i];
        cf[i].rp0 = pos;
        cf[i].rowlen = rl;
        for (int a = 0; a < nc[i]; a++) {
            rp[cf[i].mdcs + a] = pos + (int64_t)a * rl;
            lowlen[cf[i].mdcs + a] = diagoff[i] + a + 1;
            numh_ += diagoff[i] + a + 1;
        }
        pos += (int64_t)nc[i] * rl;
    }
    rp[ncsf] = pos;
    PCI_REQUIRE(pos < ((int64_t)1 << 40), PCI_E_UNSUPPORTED, "H in the CSF basis is too large");
    PCI_REQUIRE(npair < (int64_t)INT32_MAX, PCI_E_UNSUPPORTED, "too many configuration pairs");
    // the pairs (i >= j) in i order; block i sits in j's rows after j's lower
    // blocks and the upper ones of configurations before i
    std::vector<CsfPair> pairs;
    pairs.reserve(npair);
    std::vector<int64_t> run(ncf, 0);
    for (int i = 0; i < ncf; i++) {
        int64_t off = 0;
        for (int j : low[i]) {
            const int64_t offji = (j == i) ? off : lowsum[j] + run[j];
            if (j != i) run[j] += nc[i];
            pairs.push_back(CsfPair{i, j, (int)off, (int)offji});
            off += nc[j];
            ncand += (int64_t)cf[i].ndc * cf[j].ndc;
        }
    }
    Matrix &M = c.H;
    M.valid = false;
    M.has32 = false;
    M.val32.release();
    M.reset_tuning();
    upload(M.rp, rp.data(), ncsf + 1, s);
    upload(M.lowlen, lowlen.data(), ncsf, s);
    M.col.alloc((size_t)std::max<int64_t>(pos, 1));
    M.val.alloc((size_t)std::max<int64_t>(pos, 1));
    DBuf<CsfConf> dconf;
    DBuf<CsfPair> dpairs;
    upload(dconf, cf.data(), cf.size(), s);
    upload(dpairs, pairs.data(), pairs.size(), s);
    PCI_HIP(hipMemsetAsync(c.err.p, 0, 4 * sizeof(int), s));
    Tables T = c.tables();
    T.gj = 0.0;  // Hmatrix has no Gj term when kCSF /= 0 (src/csf.f90:1023, 1045)
    if (!pairs.empty()) {
        PCI_CSF_W(launch_h, c, T, (int)pairs.size(), dpairs.p, dconf.p, B.ccj.p, M.col.p, M.val.p);
        PCI_HIP(hipGetLastError());
    }
    int herr = 0;
    PCI_HIP(hipMemcpyAsync(&herr, c.err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    // numzero, Hmin over the lower triangle
    const int nb = std::max(1, std::min(1024, (ncsf + 255) / 256));
    DBuf<unsigned long long> acc;
    DBuf<double> bmin;
    acc.alloc(1);
    bmin.alloc(nb);
    PCI_HIP(hipMemsetAsync(acc.p, 0, sizeof(unsigned long long), s));
    k_csf_stats<<<nb, 256, 0, s>>>(ncsf, M.rp.p, M.lowlen.p, M.val.p, acc.p, bmin.p);
    PCI_HIP(hipGetLastError());
    unsigned long long nz = 0;
    std::vector<double> mins(nb);
    PCI_HIP(hipMemcpyAsync(&nz, acc.p, sizeof nz, hipMemcpyDeviceToHost, s));
    PCI_HIP(hipMemcpyAsync(mins.data(), bmin.p, nb * sizeof(double), hipMemcpyDeviceToHost, s));
    PCI_HIP(hipStreamSynchronize(s));
    PCI_REQUIRE(herr == 0, PCI_E_INTEGRAL, "Gint: integral is absent from CONF.INT");
    double mn = DBL_MAX;
    for (double v : mins) mn = std::min(mn, v);
    M.nd = ncsf;
    M.nnz_full = pos;
    M.numh = numh_;
    M.numzero = (int64_t)nz;
    M.hmin = numh_ > 0 ? mn : 0.0;
    M.ncand = ncand;
    M.has_res = false;  // the residual lists are the determinants', not the CSFs'
    M.csf = true;
    M.valid = true;
    M.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    c.J.valid = false;  // J^2 of the determinant basis does not apply
    c.Jf.valid = false;
    if (numh) *numh = M.numh;
    if (numzero) *numzero = M.numzero;
    if (hmin) *hmin = M.hmin;
    API_END
}

extern "C" int pci_gpu_csf_unsym(pci_ctx *ctx, int32_t nvec, const double *x, double *y) {
    API_BEGIN
    PCI_REQUIRE(ctx && x && y && nvec >= 1, PCI_E_ARG, "pci_gpu_csf_unsym: bad argument");
    Ctx &c = ctx->c;
    PCI_HIP(hipSetDevice(c.device));
    PCI_REQUIRE(c.csf.valid && c.csf.ncsf > 0, PCI_E_ARG, "pci_gpu_csf_unsym: no CSF basis (pci_gpu_csf_basis)");
    hipStream_t s = c.stream;
    const int nd = c.nd, ncsf = c.csf.ncsf;
    std::vector<CsfConf> cf = conf_records(c);
    DBuf<CsfConf> dconf;
    DBuf<double> dx, dy;
    upload(dconf, cf.data(), cf.size(), s);
    upload(dx, x, (size_t)nvec * ncsf, s);
    dy.alloc((size_t)nvec * nd);
    const int64_t total = (int64_t)nvec * nd;
    k_csf_unsym<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(nd, nvec, ncsf, c.det_conf.p, dconf.p,
                                                                c.csf.ccj.p, dx.p, dy.p);
    PCI_HIP(hipGetLastError());
    PCI_HIP(hipMemcpyAsync(y, dy.p, total * sizeof(double), hipMemcpyDeviceToHost, s));
    PCI_HIP(hipStreamSynchronize(s));
    API_END
}
