// Hamiltonian and J^2 construction on the device.
//
// Reference: FormH npes=1 (src/pconf.F90:1226-1356, 1597-1627),
// RedistributeHamCSR (src/matrix_io.f90:260-502), FormJ npes=1
// (src/formj2.f90:120-184).
//
// The reference walks, for each row n, the configuration blocks ic with
// k <= n, screens a block with CompD and compares every determinant of a
// surviving block with Rspq.  For three or more electrons almost all of
// those pairs differ in three spin-orbitals.  Here one wavefront owns one
// row and walks only the residual lists of its determinant: the C(Ne,2)
// lists of determinants that share one of its (Ne-2)-position subsets.
// Every determinant coupled to the row is in at least one of them; it is
// taken from exactly one list (the one whose residual is the Ne-2 lowest
// common positions), so no pair is visited twice.  The 64 lanes handle 64
// list entries at a time on bit-packed occupations.
//
// Three passes build the CSR without knowing its size in advance:
//   1. count  : candidates per row                          -> prefix sum
//   2. fill   : matrix elements of the candidates; nonzero ones are compacted
//               with a wavefront ballot into a candidate-sized staging buffer
//   3. merge + squeeze: each row's staging segment (one ascending run per
//               residual list) is merged by column and written to its final
//               place (prefix sum of the nonzero counts)
// The matrix kept on the device holds both triangles (rows of the full
// symmetric H, columns ascending); element (n, k) for k > n is evaluated as
// <det k|H|det n>, bit-identical to the stored lower element (k, n), so the
// lower triangle pconf stores is the prefix col <= row of each row.
#include <cfloat>
#include <chrono>

#include "internal.h"

namespace pci {

struct ResIndex {
    const int64_t *ptr;
    const int *det;
    const uint64_t *key;
    const int *det_res;
    int nres, rsize;
};

template <int W, bool FILL>
__global__ __launch_bounds__(256) void k_hpairs(Tables T, const uint64_t *__restrict__ occ, int nd, ResIndex RI,
                                                int64_t *__restrict__ cnt_out, const int64_t *__restrict__ off,
                                                int *__restrict__ col_out, double *__restrict__ val_out,
                                                int64_t *__restrict__ low_out, int64_t *__restrict__ zero_out,
                                                double *__restrict__ min_out, int *__restrict__ runlen) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (row >= nd) return;  // uniform per wavefront
    uint64_t A[W];
#pragma unroll
    for (int w = 0; w < W; w++) A[w] = occ[(size_t)w * nd + row];
    int64_t cnt = 0, low = 0;
    int zeros = 0;
    double mn = DBL_MAX;
    const int64_t base = FILL ? off[row] : 0;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (int r = 0; r < RI.nres; r++) {
        const int64_t cnt_before = cnt;
        const int lid = RI.det_res[(size_t)row * RI.nres + r];
        const uint64_t key = RI.key[lid];
        // residual positions as a bit mask, and the bits below its highest position
        uint64_t rmask[W], below[W];
        int hi = -1;
#pragma unroll
        for (int w = 0; w < W; w++) rmask[w] = 0ull;
        for (int q = 0; q < RI.rsize; q++) {
            const int p = (int)((key >> (10 * q)) & 1023ull);
#pragma unroll
            for (int w = 0; w < W; w++)
                if (w == (p >> 6)) rmask[w] |= 1ull << (p & 63);
            hi = p > hi ? p : hi;
        }
#pragma unroll
        for (int w = 0; w < W; w++) {
            const int hw = hi >> 6;
            below[w] = hi < 0 ? 0ull : (w < hw ? ~0ull : (w == hw ? ((1ull << (hi & 63)) - 1ull) : 0ull));
        }
        const int64_t j0 = RI.ptr[lid], j1 = RI.ptr[lid + 1];
        for (int64_t jb = j0; jb < j1; jb += 64) {
            const int64_t j = jb + lane;
            const bool valid = j < j1;
            const int k = valid ? RI.det[j] : 0;
            uint64_t B[W];
            int pc = 0;
            bool canon = true;
#pragma unroll
            for (int w = 0; w < W; w++) {
                B[w] = valid ? occ[(size_t)w * nd + k] : 0ull;
                pc += popc64(A[w] ^ B[w]);
                if ((A[w] & B[w] & ~rmask[w]) & below[w]) canon = false;
            }
            const bool cand = valid && canon;
            if (!FILL) {
                cnt += __popcll(__ballot(cand));
                continue;
            }
            double v = 0.0;
            bool nz = false;
            if (cand) {
                const int nf = pc >> 1;
                // the lower-triangle element (max, min) with E_k = Diag(min)
                // for an energy-dependent sigma (src/pconf.F90:1275)
                const double ek = T.kdsig ? T.dg[k <= row ? k : row] : 0.0;
                v = (k <= row) ? element<W>(T, A, B, nf, ek) : element<W>(T, B, A, nf, ek);
                nz = (v != 0.0);
                if (k <= row) {
                    if (!nz) zeros++;
                    else mn = v < mn ? v : mn;
                }
            }
            const uint64_t bal = __ballot(nz);
            if (nz) {
                const int64_t pos = base + cnt + __popcll(bal & lt);
                col_out[pos] = k;
                val_out[pos] = v;
            }
            cnt += __popcll(bal);
            low += __popcll(__ballot(nz && k <= row));
        }
        if (FILL && lane == 0) runlen[(size_t)row * RI.nres + r] = (int)(cnt - cnt_before);
    }
    if (FILL) {
        for (int o = 32; o > 0; o >>= 1) {
            zeros += __shfl_xor(zeros, o);
            const double t = __shfl_xor(mn, o);
            mn = t < mn ? t : mn;
        }
    }
    if (lane == 0) {
        cnt_out[row] = cnt;
        if (FILL) {
            low_out[row] = low;
            zero_out[row] = zeros;
            min_out[row] = mn;
        }
    }
}

// J^2 (FormJ): pairs inside the row's own configuration block, k <= n,
// differing in 0 or 2 spin-orbitals; ascending columns by construction.
template <int W, bool FILL>
__global__ __launch_bounds__(256) void k_jpairs(Tables T, const uint64_t *__restrict__ occ, int nd,
                                                const int *__restrict__ det_conf, const int *__restrict__ conf_start,
                                                int64_t *__restrict__ cnt_out, const int64_t *__restrict__ off,
                                                int *__restrict__ col_out, double *__restrict__ val_out) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (row >= nd) return;
    uint64_t A[W];
#pragma unroll
    for (int w = 0; w < W; w++) A[w] = occ[(size_t)w * nd + row];
    const int c1 = det_conf[row];
    const int k0 = conf_start[c1], k1 = row + 1;
    int64_t cnt = 0;
    const int64_t base = FILL ? off[row] : 0;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (int kb = k0; kb < k1; kb += 64) {
        const int k = kb + lane;
        const bool valid = k < k1;
        uint64_t B[W];
        int pc = 0;
#pragma unroll
        for (int w = 0; w < W; w++) {
            B[w] = valid ? occ[(size_t)w * nd + k] : 0ull;
            pc += popc64(A[w] ^ B[w]);
        }
        const bool cand = valid && (pc == 0 || pc == 4);
        if (!FILL) {
            cnt += __popcll(__ballot(cand));
            continue;
        }
        double v = 0.0;
        if (cand) v = j2_element<W>(T, A, B, pc >> 1);
        const bool nz = cand && v != 0.0;
        const uint64_t bal = __ballot(nz);
        if (nz) {
            const int64_t pos = base + cnt + __popcll(bal & lt);
            col_out[pos] = k;
            val_out[pos] = v;
        }
        cnt += __popcll(bal);
    }
    if (lane == 0) cnt_out[row] = cnt;
}

// A row's staging segment is nres runs, each ascending in column (one per
// residual list) and disjoint.  Merge them into the final row: the final
// position of an element is its index in its own run plus, for every other
// run, the number of its elements with a smaller column (binary search).
// One wavefront per row; no size limit, deterministic.
constexpr int MAX_RES = 28;  // C(8, 2)
__global__ __launch_bounds__(256) void k_merge_rows(int nd, int nres, const int64_t *__restrict__ coff,
                                                    const int64_t *__restrict__ rp, const int *__restrict__ runlen,
                                                    const int *__restrict__ ccol, const double *__restrict__ cval,
                                                    int *__restrict__ col, double *__restrict__ val) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= nd) return;
    const int64_t src = coff[row], dst = rp[row];
    const int64_t len = rp[row + 1] - rp[row];
    int off[MAX_RES + 1];
    off[0] = 0;
    for (int r = 0; r < nres; r++) off[r + 1] = off[r] + runlen[(size_t)row * nres + r];
    const int *seg = ccol + src;
    for (int64_t e = lane; e < len; e += 64) {
        int me = 0;
        while (me + 1 < nres && e >= off[me + 1]) me++;
        const int key = seg[e];
        int64_t rank = e - off[me];
        for (int r = 0; r < nres; r++) {
            if (r == me) continue;
            int lo = off[r], hi = off[r + 1];  // first index in [lo, hi) with seg >= key
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (seg[mid] < key) lo = mid + 1; else hi = mid;
            }
            rank += lo - off[r];
        }
        col[dst + rank] = key;
        val[dst + rank] = cval[src + e];
    }
}

// ---------------------------------------------------------------------------
// Device-wide exclusive scan of int64 (out has n+1 entries, out[n] = total).
// ---------------------------------------------------------------------------
constexpr int SCAN_BS = 1024;

__global__ __launch_bounds__(SCAN_BS) void k_scan_block(const int64_t *__restrict__ in, int64_t *__restrict__ out,
                                                        int64_t n, int64_t *__restrict__ bsum) {
    __shared__ int64_t s[SCAN_BS];
    const int64_t i = (int64_t)blockIdx.x * SCAN_BS + threadIdx.x;
    const int64_t x = i < n ? in[i] : 0;
    s[threadIdx.x] = x;
    __syncthreads();
    for (int off = 1; off < SCAN_BS; off <<= 1) {
        const int64_t v = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += v;
        __syncthreads();
    }
    if (i < n) out[i] = s[threadIdx.x] - x;  // exclusive
    if (threadIdx.x == SCAN_BS - 1) bsum[blockIdx.x] = s[SCAN_BS - 1];
}

__global__ void k_scan_add(int64_t *__restrict__ out, int64_t n, const int64_t *__restrict__ boff) {
    const int64_t i = (int64_t)blockIdx.x * SCAN_BS + threadIdx.x;
    if (i < n) out[i] += boff[blockIdx.x];
}

__global__ void k_set_total(int64_t *out, int64_t n, const int64_t *bsum_scanned, int64_t nb) {
    // out[n] = exclusive-scan total = scanned block offsets at nb
    out[n] = bsum_scanned[nb];
}

void scan_exclusive(const int64_t *in, int64_t *out, int64_t n, hipStream_t s, DBuf<char> &work) {
    const int64_t nb = (n + SCAN_BS - 1) / SCAN_BS;
    DBuf<int64_t> bsum, boff;
    bsum.alloc(nb + 1);
    boff.alloc(nb + 1);
    if (n > 0) {
        k_scan_block<<<(unsigned)nb, SCAN_BS, 0, s>>>(in, out, n, bsum.p);
        PCI_HIP(hipGetLastError());
    }
    if (nb > 1) {
        scan_exclusive(bsum.p, boff.p, nb, s, work);
        k_scan_add<<<(unsigned)nb, SCAN_BS, 0, s>>>(out, n, boff.p);
        PCI_HIP(hipGetLastError());
        k_set_total<<<1, 1, 0, s>>>(out, n, boff.p, nb);
    } else {
        // single block: total = bsum[0] (or 0 when n == 0)
        if (n == 0) PCI_HIP(hipMemsetAsync(out, 0, sizeof(int64_t), s));
        else PCI_HIP(hipMemcpyAsync(out + n, bsum.p, sizeof(int64_t), hipMemcpyDeviceToDevice, s));
    }
    PCI_HIP(hipGetLastError());
    PCI_HIP(hipStreamSynchronize(s));  // bsum/boff are freed on return
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_squeeze(int nd, const int64_t *__restrict__ coff,
                                                 const int64_t *__restrict__ rp, const int *__restrict__ ccol,
                                                 const double *__restrict__ cval, int *__restrict__ col,
                                                 double *__restrict__ val) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= nd) return;
    const int64_t src = coff[row], dst = rp[row], len = rp[row + 1] - rp[row];
    for (int64_t j = lane; j < len; j += 64) {
        col[dst + j] = ccol[src + j];
        val[dst + j] = cval[src + j];
    }
}

// Block reduction of the per-row statistics.
__global__ __launch_bounds__(256) void k_stats(int nd, const int64_t *__restrict__ low,
                                               const int64_t *__restrict__ zeros, const double *__restrict__ mins,
                                               unsigned long long *__restrict__ acc, double *__restrict__ bmin) {
    __shared__ int64_t s_low[256], s_z[256];
    __shared__ double s_m[256];
    int64_t l = 0, z = 0;
    double m = DBL_MAX;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < nd; i += (int64_t)gridDim.x * 256) {
        l += low[i];
        z += zeros[i];
        m = mins[i] < m ? mins[i] : m;
    }
    s_low[threadIdx.x] = l; s_z[threadIdx.x] = z; s_m[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            s_low[threadIdx.x] += s_low[threadIdx.x + o];
            s_z[threadIdx.x] += s_z[threadIdx.x + o];
            s_m[threadIdx.x] = s_m[threadIdx.x + o] < s_m[threadIdx.x] ? s_m[threadIdx.x + o] : s_m[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(&acc[0], (unsigned long long)s_low[0]);
        atomicAdd(&acc[1], (unsigned long long)s_z[0]);
        bmin[blockIdx.x] = s_m[0];
    }
}

template <int W>
static void launch_pairs(Ctx &c, bool j2, bool fill, int64_t *cnt, const int64_t *off, int *col, double *val,
                         int64_t *low, int64_t *zero, double *mn, int *runlen) {
    const Tables T = c.tables();
    const unsigned grid = (unsigned)((c.nd + 3) / 4);
    if (j2) {
        if (fill)
            k_jpairs<W, true><<<grid, 256, 0, c.stream>>>(T, c.occ.p, c.nd, c.det_conf.p, c.conf_start.p, cnt, off,
                                                          col, val);
        else
            k_jpairs<W, false><<<grid, 256, 0, c.stream>>>(T, c.occ.p, c.nd, c.det_conf.p, c.conf_start.p, cnt, off,
                                                           col, val);
    } else {
        const ResIndex RI{c.list_ptr.p, c.list_det.p, c.list_key.p, c.det_res.p, c.nres, c.rsize};
        if (fill)
            k_hpairs<W, true><<<grid, 256, 0, c.stream>>>(T, c.occ.p, c.nd, RI, cnt, off, col, val, low, zero, mn,
                                                          runlen);
        else
            k_hpairs<W, false><<<grid, 256, 0, c.stream>>>(T, c.occ.p, c.nd, RI, cnt, off, col, val, low, zero, mn,
                                                           runlen);
    }
    PCI_HIP(hipGetLastError());
}

static void pairs_dispatch(Ctx &c, bool j2, bool fill, int64_t *cnt, const int64_t *off, int *col, double *val,
                           int64_t *low, int64_t *zero, double *mn, int *runlen = nullptr) {
    switch (c.W) {
#define PCI_W(w) case w: launch_pairs<w>(c, j2, fill, cnt, off, col, val, low, zero, mn, runlen); break;
        PCI_W(1) PCI_W(2) PCI_W(3) PCI_W(4) PCI_W(5) PCI_W(6) PCI_W(7) PCI_W(8)
        PCI_W(9) PCI_W(10) PCI_W(11) PCI_W(12) PCI_W(13) PCI_W(14) PCI_W(15) PCI_W(16)
#undef PCI_W
        default: throw Error(PCI_E_UNSUPPORTED, "unsupported word count");
    }
}

// J^2 rows have no lower-part statistics to gather
__global__ void k_jstats(int nd, const int64_t *__restrict__ cnt, int64_t *__restrict__ low,
                         int64_t *__restrict__ zero, double *__restrict__ mn) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= nd) return;
    low[row] = cnt[row];
    zero[row] = 0;
    mn[row] = DBL_MAX;
}

// The space's residual lists as seen by a matrix built on it (for its
// residual-block SpMM copy); off for matrices that are not H of this space.
static void attach_res(Ctx &c, Matrix &M, bool on) {
    M.has_res = on && c.have_space && c.nlists > 0;
    if (!M.has_res) return;
    ResView &R = M.res;
    R.list_ptr = c.list_ptr.p;
    R.list_det = c.list_det.p;
    R.det_res = c.det_res.p;
    R.occ = c.occ.p;
    R.nlists = c.nlists;
    R.nres = c.nres;
    R.rsize = c.rsize;
    R.ne = c.ne;
    R.W = c.W;
}

void build_matrix(Ctx &c, bool j2, Matrix &M) {
    PCI_REQUIRE(c.have_space && c.have_integrals, PCI_E_ARG, "set the space and the integrals first");
    hipStream_t s = c.stream;
    const int nd = c.nd;
    M.valid = false;
    M.has32 = false;
    M.val32.release();
    DBuf<int64_t> ccnt, coff, low, zero;
    DBuf<double> mins;
    ccnt.alloc(nd);
    coff.alloc(nd + 1);
    PCI_HIP(hipMemsetAsync(c.err.p, 0, 4 * sizeof(int), s));
    // 1. candidates
    pairs_dispatch(c, j2, false, ccnt.p, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    scan_exclusive(ccnt.p, coff.p, nd, s, c.work);
    int64_t ncand = 0;
    PCI_HIP(hipMemcpyAsync(&ncand, coff.p + nd, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    PCI_HIP(hipStreamSynchronize(s));
    // 2. values into the staging buffer
    DBuf<int> ccol;
    DBuf<double> cval;
    ccol.alloc(ncand);
    cval.alloc(ncand);
    low.alloc(nd);
    zero.alloc(nd);
    mins.alloc(nd);
    DBuf<int> runlen;
    PCI_REQUIRE(c.nres <= MAX_RES, PCI_E_UNSUPPORTED, "too many residual lists per determinant");
    if (!j2) runlen.alloc((size_t)nd * c.nres);
    pairs_dispatch(c, j2, true, ccnt.p, coff.p, ccol.p, cval.p, low.p, zero.p, mins.p, runlen.p);
    if (j2) k_jstats<<<(nd + 255) / 256, 256, 0, s>>>(nd, ccnt.p, low.p, zero.p, mins.p);
    int herr = 0;
    PCI_HIP(hipMemcpyAsync(&herr, c.err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    PCI_HIP(hipStreamSynchronize(s));
    PCI_REQUIRE(herr == 0, PCI_E_INTEGRAL, "Gint: integral is absent from CONF.INT");
    // 3. final CSR
    M.rp.alloc(nd + 1);
    scan_exclusive(ccnt.p, M.rp.p, nd, s, c.work);
    int64_t nnz = 0;
    PCI_HIP(hipMemcpyAsync(&nnz, M.rp.p + nd, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    PCI_HIP(hipStreamSynchronize(s));
    M.col.alloc(nnz);
    M.val.alloc(nnz);
    if (j2) {
        k_squeeze<<<(nd + 3) / 4, 256, 0, s>>>(nd, coff.p, M.rp.p, ccol.p, cval.p, M.col.p, M.val.p);
    } else {
        k_merge_rows<<<(nd + 3) / 4, 256, 0, s>>>(nd, c.nres, coff.p, M.rp.p, runlen.p, ccol.p, cval.p, M.col.p,
                                                  M.val.p);
    }
    PCI_HIP(hipGetLastError());
    // statistics: NumH, numzero, Hmin
    const int sb = 256;
    DBuf<unsigned long long> acc;
    DBuf<double> bmin;
    acc.alloc(2);
    bmin.alloc(sb);
    PCI_HIP(hipMemsetAsync(acc.p, 0, 2 * sizeof(unsigned long long), s));
    k_stats<<<sb, 256, 0, s>>>(nd, low.p, zero.p, mins.p, acc.p, bmin.p);
    PCI_HIP(hipGetLastError());
    unsigned long long ha[2];
    std::vector<double> hb(sb);
    PCI_HIP(hipMemcpyAsync(ha, acc.p, sizeof(ha), hipMemcpyDeviceToHost, s));
    PCI_HIP(hipMemcpyAsync(hb.data(), bmin.p, sb * sizeof(double), hipMemcpyDeviceToHost, s));
    M.lowlen.alloc(nd);
    PCI_HIP(hipMemcpyAsync(M.lowlen.p, low.p, nd * sizeof(int64_t), hipMemcpyDeviceToDevice, s));
    PCI_HIP(hipStreamSynchronize(s));
    double hmin = DBL_MAX;
    for (double v : hb) hmin = v < hmin ? v : hmin;
    M.nd = nd;
    M.nnz_full = nnz;
    M.numh = (int64_t)ha[0];
    M.numzero = (int64_t)ha[1];
    M.ncand = ncand;
    M.hmin = M.numh > 0 ? hmin : 0.0;
    M.reset_tuning();
    attach_res(c, M, !j2);
    M.csf = false;
    M.valid = true;
}

// ---------------------------------------------------------------------------
// H built straight into the residual blocks (no CSR).  A workgroup takes one
// unit of the block layout (a list and a range of its 16-row chunks) and
// stages the occupations of the list's determinants in LDS; each wavefront
// takes a chunk and walks its column steps, lane 16 j + (p mod 16) computing
// the two values (p, 8 s + 4 h + j), h = 0, 1, that the SpMM reads as one
// pair, so every tile of the block is written once, as one coalesced store.
// A pair is evaluated where this list is its canonical one (the Ne-2 lowest
// positions the two share: the test k_hpairs uses), as the lower-triangle
// element (max, min) like k_hpairs, so the block is exactly symmetric and
// holds, value for value, what the CSR build stores; every other position is
// 0.  Per wavefront: pairs evaluated, nonzero values, and over the lower
// triangle (q <= p) the stored count, the exact zeros and the minimum.
struct RbStats {
    int64_t cand, nnz, low, zero;
    double mn;
};

template <int W>
__global__ __launch_bounds__(256) void k_rb_build(Tables T, const uint64_t *__restrict__ occ, int nd,
                                                  const int *__restrict__ unit, const int64_t *__restrict__ list_ptr,
                                                  const int *__restrict__ list_det, const uint64_t *__restrict__ key,
                                                  int rsize, const int64_t *__restrict__ boff,
                                                  double *__restrict__ out, RbStats *__restrict__ st) {
    extern __shared__ __align__(16) char rbb_lds[];
    uint64_t *sOcc = (uint64_t *)rbb_lds;  // n x W
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int lid = unit[3 * blockIdx.x], ch0 = unit[3 * blockIdx.x + 1], ch1 = unit[3 * blockIdx.x + 2];
    const int64_t lp = list_ptr[lid];
    const int n = (int)(list_ptr[lid + 1] - lp);
    const int steps = (n + 7) >> 3;
    for (int e = tid; e < n * W; e += 256) {
        const int q = e / W, w = e - q * W;
        sOcc[e] = occ[(size_t)w * nd + list_det[lp + q]];
    }
    // the residual's positions as a mask, and the bits below its highest one
    uint64_t rmask[W], below[W];
    {
        const uint64_t kk = key[lid];
        int hi = -1;
#pragma unroll
        for (int w = 0; w < W; w++) rmask[w] = 0ull;
        for (int q = 0; q < rsize; q++) {
            const int pos = (int)((kk >> (10 * q)) & 1023ull);
#pragma unroll
            for (int w = 0; w < W; w++)
                if (w == (pos >> 6)) rmask[w] |= 1ull << (pos & 63);
            hi = pos > hi ? pos : hi;
        }
#pragma unroll
        for (int w = 0; w < W; w++) {
            const int hw = hi >> 6;
            below[w] = hi < 0 ? 0ull : (w < hw ? ~0ull : (w == hw ? ((1ull << (hi & 63)) - 1ull) : 0ull));
        }
    }
    __syncthreads();
    int64_t cand = 0, nnz = 0, low = 0, zero = 0;
    double mn = DBL_MAX;
    const int sub = lane >> 4;  // the lane's column in each half of a step
    for (int c = ch0 + wv; c < ch1; c += 4) {
        const int p = c * RB_R + (lane & 15);
        const bool prow = p < n;
        uint64_t A[W];
#pragma unroll
        for (int w = 0; w < W; w++) A[w] = prow ? sOcc[p * W + w] : 0ull;
        const int kp = prow ? list_det[lp + p] : 0;
        double2 *dst = (double2 *)(out + boff[lid]) + (size_t)c * steps * 64 + lane;
        for (int s = 0; s < steps; s++) {
            double v2[2];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int q = 8 * s + 4 * h + sub;
                double v = 0.0;
                if (prow && q < n) {
                    uint64_t B[W];
                    bool canon = true;
                    int pc = 0;
#pragma unroll
                    for (int w = 0; w < W; w++) {
                        B[w] = sOcc[q * W + w];
                        pc += popc64(A[w] ^ B[w]);
                        if ((A[w] & B[w] & ~rmask[w]) & below[w]) canon = false;
                    }
                    if (canon) {
                        const int kq = list_det[lp + q];
                        const double ek = T.kdsig ? T.dg[kq <= kp ? kq : kp] : 0.0;
                        v = (kq <= kp) ? element<W>(T, A, B, pc >> 1, ek) : element<W>(T, B, A, pc >> 1, ek);
                        cand++;
                        if (v != 0.0) nnz++;
                        if (q <= p) {
                            if (v == 0.0) zero++;
                            else {
                                low++;
                                mn = v < mn ? v : mn;
                            }
                        }
                    }
                }
                v2[h] = v;
            }
            dst[(size_t)s * 64] = make_double2(v2[0], v2[1]);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        cand += __shfl_xor(cand, o);
        nnz += __shfl_xor(nnz, o);
        low += __shfl_xor(low, o);
        zero += __shfl_xor(zero, o);
        const double t = __shfl_xor(mn, o);
        mn = t < mn ? t : mn;
    }
    if (lane == 0) st[(size_t)blockIdx.x * 4 + wv] = RbStats{cand, nnz, low, zero, mn};
}

// FP32 blocks: every value rounded to FP32, then each determinant's diagonal
// (its first list, its own position) replaced by the rounding of H_ii - hmin
__global__ __launch_bounds__(256) void k_rb_to32(int64_t n, const double *__restrict__ v64, float *__restrict__ v32) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        v32[i] = (float)v64[i];
}

__global__ __launch_bounds__(256) void k_rb_diag32(int nd, ResView R, const int *__restrict__ det_pos,
                                                   const int64_t *__restrict__ boff, const double *__restrict__ v64,
                                                   double hmin, float *__restrict__ v32) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nd) return;
    const int lid = R.det_res[(size_t)i * R.nres], p = det_pos[(size_t)i * R.nres];
    const int n = (int)(R.list_ptr[lid + 1] - R.list_ptr[lid]);
    const int64_t at = boff[lid] + rb_index(p, p, (n + 7) >> 3);
    v32[at] = (float)(v64[at] - hmin);
}

template <int W>
static void rb_build_launch(Ctx &c, const RblkCopy &B, int maxn, double *out, RbStats *st) {
    const size_t shm = (size_t)maxn * W * sizeof(uint64_t);
    k_rb_build<W><<<(unsigned)B.nunit, 256, shm, c.stream>>>(c.tables(), c.occ.p, c.nd, B.unit.p, c.list_ptr.p,
                                                              c.list_det.p, c.list_key.p, c.rsize, B.boff.p, out, st);
    PCI_HIP(hipGetLastError());
}

static void build_blocks(Ctx &c, Matrix &M) {
    PCI_REQUIRE(c.have_space && c.have_integrals, PCI_E_ARG, "set the space and the integrals first");
    PCI_REQUIRE(c.nlists > 0 && c.ne <= 8, PCI_E_UNSUPPORTED,
                "residual-block build: needs the residual lists (2 to 8 valence electrons)");
    hipStream_t s = c.stream;
    const int nd = c.nd;
    // drop whatever H was resident before
    M.valid = false;
    M.has32 = false;
    M.blocks_only = false;
    M.val32.release();
    M.rp.release();
    M.col.release();
    M.val.release();
    M.lowlen.release();
    M.reset_tuning();
    M.nd = nd;
    attach_res(c, M, true);
    RblkCopy &B64 = M.rblk[1], &B32 = M.rblk[0];
    const int64_t nvals = rblk_layout(M, B64, s);
    (void)rblk_layout(M, B32, s);
    const size_t lds = (size_t)B64.maxn * c.W * sizeof(uint64_t);
    PCI_REQUIRE(lds <= 64 * 1024, PCI_E_UNSUPPORTED,
                "residual-block build: a list of " + std::to_string(B64.maxn) + " determinants exceeds the LDS");
    {
        size_t fr = 0, tot = 0;
        PCI_HIP(hipMemGetInfo(&fr, &tot));
        const size_t need = (size_t)nvals * 12 + (size_t)nd * c.nres * 8 + (size_t)B64.nunit * 4 * sizeof(RbStats);
        PCI_REQUIRE(need + ((size_t)2 << 30) <= fr, PCI_E_NOMEM,
                    "residual-block H of " + std::to_string(need) + " bytes does not fit the free device memory (" +
                        std::to_string(fr) + " bytes)");
    }
    rblk_positions(M, s);
    B64.val.alloc((size_t)nvals * sizeof(double));
    DBuf<RbStats> st;
    st.alloc((size_t)B64.nunit * 4);
    PCI_HIP(hipMemsetAsync(st.p, 0, st.bytes(), s));
    PCI_HIP(hipMemsetAsync(c.err.p, 0, 4 * sizeof(int), s));
    switch (c.W) {
#define PCI_W(w) case w: rb_build_launch<w>(c, B64, B64.maxn, (double *)B64.val.p, st.p); break;
        PCI_W(1) PCI_W(2) PCI_W(3) PCI_W(4) PCI_W(5) PCI_W(6) PCI_W(7) PCI_W(8)
        PCI_W(9) PCI_W(10) PCI_W(11) PCI_W(12) PCI_W(13) PCI_W(14) PCI_W(15) PCI_W(16)
#undef PCI_W
        default: throw Error(PCI_E_UNSUPPORTED, "unsupported word count");
    }
    std::vector<RbStats> hs((size_t)B64.nunit * 4);
    int herr = 0;
    PCI_HIP(hipMemcpyAsync(hs.data(), st.p, st.bytes(), hipMemcpyDeviceToHost, s));
    PCI_HIP(hipMemcpyAsync(&herr, c.err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    PCI_HIP(hipStreamSynchronize(s));
    PCI_REQUIRE(herr == 0, PCI_E_INTEGRAL, "Gint: integral is absent from CONF.INT");
    int64_t cand = 0, nnz = 0, low = 0, zero = 0;
    double mn = DBL_MAX;
    for (const RbStats &r : hs) {  // a wavefront without work reports zeros and mn = 0: skip its mn
        cand += r.cand;
        nnz += r.nnz;
        low += r.low;
        zero += r.zero;
        if (r.low > 0) mn = r.mn < mn ? r.mn : mn;
    }
    M.nnz_full = nnz;
    M.numh = low;
    M.numzero = zero;
    M.ncand = cand;
    M.hmin = low > 0 ? mn : 0.0;
    B64.nvals = nvals;
    B64.built = true;
    // FP32 blocks with the shifted diagonal
    B32.val.alloc((size_t)nvals * sizeof(float));
    k_rb_to32<<<4096, 256, 0, s>>>(nvals, (const double *)B64.val.p, (float *)B32.val.p);
    k_rb_diag32<<<(unsigned)((nd + 255) / 256), 256, 0, s>>>(nd, M.res, M.det_pos.p, B64.boff.p,
                                                              (const double *)B64.val.p, M.hmin, (float *)B32.val.p);
    PCI_HIP(hipGetLastError());
    PCI_HIP(hipStreamSynchronize(s));
    B32.nvals = nvals;
    B32.built = true;
    M.shift32 = 0.f;
    M.has32 = true;
    M.blocks_only = true;
    M.csf = false;
    M.valid = true;
}

// The rows of a blocks-only H: per row and list, the nonzero values of the
// row's block row with column <= row (FILL: one ascending run per list,
// compacted with a ballot, 1-based columns), merged per row by k_merge_rows.
template <bool FILL>
__global__ __launch_bounds__(256) void k_rb_rows(int row0, int nrows, ResView R, const int *__restrict__ det_pos,
                                                 const int64_t *__restrict__ boff, const double *__restrict__ b64,
                                                 int *__restrict__ runlen, const int64_t *__restrict__ off,
                                                 int *__restrict__ ocol, double *__restrict__ oval) {
    const int lane = threadIdx.x & 63;
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= nrows) return;
    const int row = row0 + t;
    const uint64_t lt = (1ull << lane) - 1ull;
    int64_t at = FILL ? off[t] : 0;
    for (int q = 0; q < R.nres; q++) {
        const int lid = R.det_res[(size_t)row * R.nres + q], p = det_pos[(size_t)row * R.nres + q];
        const int64_t lp = R.list_ptr[lid];
        const int n = (int)(R.list_ptr[lid + 1] - lp);
        const int steps = (n + 7) >> 3;
        const double *blk = b64 + boff[lid];
        int cnt = 0;
        for (int b = 0; b < n; b += 64) {
            const int j = b + lane;
            double v = 0.0;
            int k = 0;
            if (j < n) {
                k = R.list_det[lp + j];
                v = blk[rb_index(p, j, steps)];
            }
            const bool keep = j < n && v != 0.0 && k <= row;
            const uint64_t bal = __ballot(keep);
            if (FILL && keep) {
                const int64_t e = at + cnt + __popcll(bal & lt);
                ocol[e] = k + 1;
                oval[e] = v;
            }
            cnt += __popcll(bal);
        }
        if (!FILL && lane == 0) runlen[(size_t)t * R.nres + q] = cnt;
        at += cnt;
    }
}

__global__ void k_runlen_sum(int nrows, int nres, const int *__restrict__ runlen, int64_t *__restrict__ cnt) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrows) return;
    int64_t m = 0;
    for (int q = 0; q < nres; q++) m += runlen[(size_t)t * nres + q];
    cnt[t] = m;
}

void rblk_lower_rows(Ctx &c, Matrix &M, int row0, int nrows, int64_t *rowptr, int32_t *col, double *val,
                     int64_t cap) {
    hipStream_t s = c.stream;
    const RblkCopy &B = M.rblk[1];
    const unsigned g = (unsigned)((nrows + 3) / 4);
    DBuf<int> runlen;
    DBuf<int64_t> cnt, lrp;
    runlen.alloc((size_t)nrows * M.res.nres);
    cnt.alloc(nrows);
    lrp.alloc((size_t)nrows + 1);
    k_rb_rows<false><<<g, 256, 0, s>>>(row0, nrows, M.res, M.det_pos.p, B.boff.p, (const double *)B.val.p, runlen.p,
                                       nullptr, nullptr, nullptr);
    k_runlen_sum<<<(unsigned)((nrows + 255) / 256), 256, 0, s>>>(nrows, M.res.nres, runlen.p, cnt.p);
    PCI_HIP(hipGetLastError());
    scan_exclusive(cnt.p, lrp.p, nrows, s, c.work);
    PCI_HIP(hipMemcpyAsync(rowptr, lrp.p, (nrows + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    PCI_HIP(hipStreamSynchronize(s));
    const int64_t n = rowptr[nrows];
    if (!col && !val) return;
    PCI_REQUIRE(cap >= n, PCI_E_ARG, "pci_gpu_get_csr_rows: output buffers hold " + std::to_string(cap) +
                                         " elements, the rows have " + std::to_string(n));
    DBuf<int> sc, oc;
    DBuf<double> sv, ov;
    sc.alloc(n);
    sv.alloc(n);
    oc.alloc(n);
    ov.alloc(n);
    k_rb_rows<true><<<g, 256, 0, s>>>(row0, nrows, M.res, M.det_pos.p, B.boff.p, (const double *)B.val.p, nullptr,
                                      lrp.p, sc.p, sv.p);
    k_merge_rows<<<g, 256, 0, s>>>(nrows, M.res.nres, lrp.p, lrp.p, runlen.p, sc.p, sv.p, oc.p, ov.p);
    PCI_HIP(hipGetLastError());
    if (col) PCI_HIP(hipMemcpyAsync(col, oc.p, n * sizeof(int), hipMemcpyDeviceToHost, s));
    if (val) PCI_HIP(hipMemcpyAsync(val, ov.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
    PCI_HIP(hipStreamSynchronize(s));
}

// FP32 copy with the diagonal shifted by -hmin (H' = H - Hmin I, davidson.f90:630-633).
// The shift is applied in FP64 before rounding, so H'_ii carries one FP32
// rounding of the shifted value.
__global__ __launch_bounds__(256) void k_to32(int nd, const int64_t *__restrict__ rp, const int *__restrict__ col,
                                              const double *__restrict__ val, double hmin, float *__restrict__ v32) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= nd) return;
    for (int64_t j = rp[row] + lane; j < rp[row + 1]; j += 64) {
        const double v = val[j];
        v32[j] = (float)(col[j] == row ? v - hmin : v);
    }
}

// Rows whose diagonal element was dropped as an exact zero.
__global__ void k_missing_diag(int nd, const int64_t *__restrict__ rp, const int *__restrict__ col,
                               const int64_t *__restrict__ low, unsigned long long *__restrict__ cnt) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= nd) return;
    const int64_t l = low[row];
    if (l == 0 || col[rp[row] + l - 1] != row) atomicAdd(cnt, 1ull);
}

void ensure_fp32(Ctx &c) {
    Matrix &M = c.H;
    PCI_REQUIRE(M.valid, PCI_E_ARG, "no Hamiltonian on the device");
    if (M.has32) return;
    DBuf<unsigned long long> miss;
    miss.alloc(1);
    PCI_HIP(hipMemsetAsync(miss.p, 0, sizeof(unsigned long long), c.stream));
    k_missing_diag<<<(M.nd + 255) / 256, 256, 0, c.stream>>>(M.nd, M.rp.p, M.col.p, M.lowlen.p, miss.p);
    unsigned long long nmiss = 0;
    PCI_HIP(hipMemcpyAsync(&nmiss, miss.p, sizeof(nmiss), hipMemcpyDeviceToHost, c.stream));
    PCI_HIP(hipStreamSynchronize(c.stream));
    M.val32.alloc(M.nnz_full);
    // with an absent diagonal the shift is applied in the SpMM instead
    const double sh = nmiss ? 0.0 : M.hmin;
    M.shift32 = nmiss ? (float)(-M.hmin) : 0.f;
    k_to32<<<(M.nd + 3) / 4, 256, 0, c.stream>>>(M.nd, M.rp.p, M.col.p, M.val.p, sh, M.val32.p);
    PCI_HIP(hipGetLastError());
    PCI_HIP(hipStreamSynchronize(c.stream));
    M.has32 = true;
}

// Lower-triangle extraction in pCONF.HIJ form (1-based columns).
__global__ __launch_bounds__(256) void k_lower(int nd, const int64_t *__restrict__ rp, const int *__restrict__ col,
                                               const double *__restrict__ val, const int64_t *__restrict__ lrp,
                                               int *__restrict__ ocol, double *__restrict__ oval) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= nd) return;
    const int64_t src = rp[row], dst = lrp[row], len = lrp[row + 1] - lrp[row];
    for (int64_t j = lane; j < len; j += 64) {
        ocol[dst + j] = col[src + j] + 1;
        oval[dst + j] = val[src + j];
    }
}

static void copy_lower(Ctx &c, Matrix &M, int64_t *rowptr, int32_t *col, double *val) {
    PCI_REQUIRE(M.valid, PCI_E_ARG, "matrix not built");
    if (M.blocks_only) {
        std::vector<int64_t> rp(rowptr ? 0 : (size_t)M.nd + 1);
        rblk_lower_rows(c, M, 0, M.nd, rowptr ? rowptr : rp.data(), col, val, M.numh);
        return;
    }
    hipStream_t s = c.stream;
    DBuf<int64_t> lrp;
    lrp.alloc(M.nd + 1);
    scan_exclusive(M.lowlen.p, lrp.p, M.nd, s, c.work);
    DBuf<int> oc;
    DBuf<double> ov;
    oc.alloc(M.numh);
    ov.alloc(M.numh);
    k_lower<<<(M.nd + 3) / 4, 256, 0, s>>>(M.nd, M.rp.p, M.col.p, M.val.p, lrp.p, oc.p, ov.p);
    PCI_HIP(hipGetLastError());
    if (rowptr) PCI_HIP(hipMemcpyAsync(rowptr, lrp.p, (M.nd + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    if (col) PCI_HIP(hipMemcpyAsync(col, oc.p, M.numh * sizeof(int), hipMemcpyDeviceToHost, s));
    if (val) PCI_HIP(hipMemcpyAsync(val, ov.p, M.numh * sizeof(double), hipMemcpyDeviceToHost, s));
    PCI_HIP(hipStreamSynchronize(s));
}

// Rows [row0, row0 + nrows) of the lower triangle: rowptr (nrows + 1,
// starting at 0) always; columns and values when the caller's buffers (cap
// elements) are given.
static void copy_lower_rows(Ctx &c, Matrix &M, int row0, int nrows, int64_t *rowptr, int32_t *col, double *val,
                            int64_t cap) {
    PCI_REQUIRE(M.valid, PCI_E_ARG, "matrix not built");
    if (M.blocks_only) {
        rblk_lower_rows(c, M, row0, nrows, rowptr, col, val, cap);
        return;
    }
    hipStream_t s = c.stream;
    DBuf<int64_t> lrp;
    lrp.alloc((size_t)nrows + 1);
    scan_exclusive(M.lowlen.p + row0, lrp.p, nrows, s, c.work);
    PCI_HIP(hipMemcpyAsync(rowptr, lrp.p, (nrows + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    PCI_HIP(hipStreamSynchronize(s));
    const int64_t n = rowptr[nrows];
    if (!col && !val) return;
    PCI_REQUIRE(cap >= n, PCI_E_ARG, "pci_gpu_get_csr_rows: output buffers hold " + std::to_string(cap) +
                                         " elements, the rows have " + std::to_string(n));
    DBuf<int> oc;
    DBuf<double> ov;
    oc.alloc(n);
    ov.alloc(n);
    k_lower<<<(nrows + 3) / 4, 256, 0, s>>>(nrows, M.rp.p + row0, M.col.p, M.val.p, lrp.p, oc.p, ov.p);
    PCI_HIP(hipGetLastError());
    if (col) PCI_HIP(hipMemcpyAsync(col, oc.p, n * sizeof(int), hipMemcpyDeviceToHost, s));
    if (val) PCI_HIP(hipMemcpyAsync(val, ov.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
    PCI_HIP(hipStreamSynchronize(s));
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

extern "C" int pci_gpu_build_h(pci_ctx *ctx, int32_t flags, int64_t *numh, int64_t *numzero, double *hmin) {
    API_BEGIN
    PCI_REQUIRE(ctx, PCI_E_ARG, "null context");
    PCI_REQUIRE((flags & ~PCI_BUILD_BLOCKS) == 0, PCI_E_ARG, "unknown pci_gpu_build_h flag");
    Ctx &c = ctx->c;
    PCI_REQUIRE(c.ksig < 2 || c.have_sigma2, PCI_E_ARG,
                "Ksig = 2: pci_gpu_set_sigma2 (SCRC.CON and the screening coefficients) must come before the build");
    PCI_HIP(hipSetDevice(c.device));
    const auto t0 = std::chrono::steady_clock::now();
    if (flags & PCI_BUILD_BLOCKS) build_blocks(c, c.H);
    else build_matrix(c, false, c.H);
    c.H.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (numh) *numh = c.H.numh;
    if (numzero) *numzero = c.H.numzero;
    if (hmin) *hmin = c.H.hmin;
    API_END
}

extern "C" int pci_gpu_build_info(pci_ctx *ctx, int64_t *ncand, int64_t *nnz_full, double *build_ms) {
    API_BEGIN
    PCI_REQUIRE(ctx, PCI_E_ARG, "null context");
    const Matrix &M = ctx->c.H;
    PCI_REQUIRE(M.valid, PCI_E_ARG, "no Hamiltonian on the device");
    if (ncand) *ncand = M.ncand;
    if (nnz_full) *nnz_full = M.nnz_full;
    if (build_ms) *build_ms = M.build_ms;
    API_END
}

extern "C" int pci_gpu_get_csr(pci_ctx *ctx, int64_t *rowptr, int32_t *col, double *val) {
    API_BEGIN
    PCI_REQUIRE(ctx, PCI_E_ARG, "null context");
    PCI_HIP(hipSetDevice(ctx->c.device));
    copy_lower(ctx->c, ctx->c.H, rowptr, col, val);
    API_END
}

extern "C" int pci_gpu_get_csr_rows(pci_ctx *ctx, int32_t row1, int32_t nrows, int64_t *rowptr, int32_t *col,
                                    double *val, int64_t cap) {
    API_BEGIN
    PCI_REQUIRE(ctx && rowptr, PCI_E_ARG, "null argument");
    Ctx &c = ctx->c;
    PCI_HIP(hipSetDevice(c.device));
    PCI_REQUIRE(c.H.valid, PCI_E_ARG, "no Hamiltonian on the device");
    PCI_REQUIRE(row1 >= 1 && nrows >= 1 && (int64_t)row1 - 1 + nrows <= c.H.nd, PCI_E_ARG,
                "pci_gpu_get_csr_rows: rows out of range");
    copy_lower_rows(c, c.H, row1 - 1, nrows, rowptr, col, val, cap);
    API_END
}

extern "C" int pci_gpu_build_j(pci_ctx *ctx, int64_t *numj) {
    API_BEGIN
    PCI_REQUIRE(ctx, PCI_E_ARG, "null context");
    Ctx &c = ctx->c;
    PCI_HIP(hipSetDevice(c.device));
    c.Jf.valid = false;
    build_matrix(c, true, c.J);
    if (numj) *numj = c.J.numh;
    API_END
}

extern "C" int pci_gpu_get_jcsr(pci_ctx *ctx, int64_t *rowptr, int32_t *col, double *val) {
    API_BEGIN
    PCI_REQUIRE(ctx, PCI_E_ARG, "null context");
    PCI_HIP(hipSetDevice(ctx->c.device));
    copy_lower(ctx->c, ctx->c.J, rowptr, col, val);
    API_END
}

// Kl=1: build the resident full matrix from a lower-triangle CSR read from
// pCONF.HIJ (rows' columns ascending, 1-based).  The transpose is done on
// the host with a stable counting sort, so each full row is its lower part
// followed by the column's entries of later rows in ascending order.
extern "C" int pci_gpu_set_csr(pci_ctx *ctx, int32_t nd, int64_t numh, const int64_t *rowptr,
                               const int32_t *col, const double *val) {
    API_BEGIN
    PCI_REQUIRE(ctx && rowptr && col && val, PCI_E_ARG, "null argument");
    Ctx &c = ctx->c;
    PCI_HIP(hipSetDevice(c.device));
    PCI_REQUIRE(nd > 0 && rowptr[0] == 0 && rowptr[nd] == numh, PCI_E_ARG, "inconsistent CSR");
    std::vector<int64_t> up(nd + 1, 0), low(nd);
    double hmin = DBL_MAX;
    for (int n = 0; n < nd; n++) {
        PCI_REQUIRE(rowptr[n + 1] >= rowptr[n], PCI_E_ARG, "row pointer decreases");
        low[n] = rowptr[n + 1] - rowptr[n];
        int prev = 0;
        for (int64_t j = rowptr[n]; j < rowptr[n + 1]; j++) {
            PCI_REQUIRE(col[j] >= 1 && col[j] <= n + 1 && col[j] > prev, PCI_E_ARG,
                        "CSR columns must be ascending and in the lower triangle");
            prev = col[j];
            if (col[j] - 1 < n) up[col[j] - 1 + 1]++;
            hmin = val[j] < hmin ? val[j] : hmin;
        }
    }
    std::vector<int64_t> rp(nd + 1, 0);
    for (int n = 0; n < nd; n++) rp[n + 1] = rp[n] + low[n] + up[n + 1];
    const int64_t nnz = rp[nd];
    std::vector<int> fc(nnz);
    std::vector<double> fv(nnz);
    std::vector<int64_t> cur(nd);
    for (int n = 0; n < nd; n++) {
        int64_t p = rp[n];
        for (int64_t j = rowptr[n]; j < rowptr[n + 1]; j++, p++) {
            fc[p] = col[j] - 1;
            fv[p] = val[j];
        }
        cur[n] = p;
    }
    for (int n = 0; n < nd; n++)
        for (int64_t j = rowptr[n]; j < rowptr[n + 1]; j++) {
            const int k = col[j] - 1;
            if (k < n) {
                fc[cur[k]] = n;
                fv[cur[k]] = val[j];
                cur[k]++;
            }
        }
    Matrix &M = c.H;
    hipStream_t s = c.stream;
    upload(M.rp, rp.data(), nd + 1, s);
    upload(M.col, fc.data(), nnz, s);
    upload(M.val, fv.data(), nnz, s);
    upload(M.lowlen, low.data(), nd, s);
    PCI_HIP(hipStreamSynchronize(s));
    M.val32.release();
    M.has32 = false;
    M.nd = nd;
    M.nnz_full = nnz;
    M.numh = numh;
    M.numzero = 0;
    M.hmin = numh > 0 ? hmin : 0.0;
    M.reset_tuning();
    // a CSR read from file is checked against the lists when its
    // residual-block copy is built (elements outside them make it ineligible)
    attach_res(c, M, c.have_space && c.nd == nd);
    M.csf = false;
    M.valid = true;
    if (!c.have_space) c.nd = nd;
    API_END
}

namespace pci {
// J^2 with both triangles: the lower-triangle CSR is mirrored on the host
// (J^2 is block diagonal over configurations and small next to H).
void ensure_jfull(Ctx &c) {
    if (c.Jf.valid) return;
    PCI_REQUIRE(c.J.valid, PCI_E_ARG, "build J^2 (pci_gpu_build_j) first");
    hipStream_t s = c.stream;
    const int nd = c.J.nd;
    const int64_t nl = c.J.nnz_full;
    std::vector<int64_t> rp(nd + 1);
    std::vector<int> col(nl);
    std::vector<double> val(nl);
    PCI_HIP(hipMemcpyAsync(rp.data(), c.J.rp.p, (nd + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    PCI_HIP(hipMemcpyAsync(col.data(), c.J.col.p, nl * sizeof(int), hipMemcpyDeviceToHost, s));
    PCI_HIP(hipMemcpyAsync(val.data(), c.J.val.p, nl * sizeof(double), hipMemcpyDeviceToHost, s));
    PCI_HIP(hipStreamSynchronize(s));
    std::vector<int64_t> cnt(nd + 1, 0);
    for (int r = 0; r < nd; r++)
        for (int64_t j = rp[r]; j < rp[r + 1]; j++) {
            cnt[r + 1]++;
            if (col[j] != r) cnt[col[j] + 1]++;
        }
    for (int r = 0; r < nd; r++) cnt[r + 1] += cnt[r];
    const int64_t nf = cnt[nd];
    std::vector<int> fcol(nf);
    std::vector<double> fval(nf);
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    // visiting rows ascending and appending keeps every row's columns
    // ascending: row r gets its own entries (k <= r) when r is visited and
    // the mirrored ones (k > r) later, in the order of k.
    for (int r = 0; r < nd; r++)
        for (int64_t j = rp[r]; j < rp[r + 1]; j++) {
            const int k = col[j];
            fcol[pos[r]] = k;
            fval[pos[r]++] = val[j];
            if (k != r) {
                fcol[pos[k]] = r;
                fval[pos[k]++] = val[j];
            }
        }
    Matrix &M = c.Jf;
    upload(M.rp, cnt.data(), nd + 1, s);
    upload(M.col, fcol.data(), nf, s);
    upload(M.val, fval.data(), nf, s);
    PCI_HIP(hipStreamSynchronize(s));
    M.nd = nd;
    M.nnz_full = nf;
    M.numh = nl;
    M.hmin = 0.0;
    M.reset_tuning();
    attach_res(c, M, false);
    M.csf = false;
    M.valid = true;
}
}  // namespace pci

extern "C" int pci_gpu_set_jterms(pci_ctx *ctx, double xj_av, int32_t njd, const int32_t *jt) {
    API_BEGIN
    PCI_REQUIRE(ctx && (njd == 0 || jt) && njd >= 0 && njd <= 256, PCI_E_ARG, "bad J term list");
    for (int i = 0; i < njd; i++) PCI_REQUIRE(jt[i] >= 0 && (i == 0 || jt[i] > jt[i - 1]), PCI_E_ARG,
                                               "J terms must be ascending doubled J values");
    ctx->c.xj_av = xj_av;
    ctx->c.jt.assign(jt, jt + njd);
    API_END
}
