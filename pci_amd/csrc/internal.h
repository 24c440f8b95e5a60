// Internal state of a pci_gpu context and helpers shared by the .hip units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include <condition_variable>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pci_gpu.h"
#include "slater.h"

namespace pci {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);

#define PCI_HIP(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            throw ::pci::Error(e_ == hipErrorOutOfMemory ? PCI_E_NOMEM : PCI_E_HIP,         \
                               std::string(#expr) + ": " + hipGetErrorString(e_));          \
    } while (0)

#define PCI_REQUIRE(cond, code, msg)                                                        \
    do {                                                                                    \
        if (!(cond)) throw ::pci::Error((code), (msg));                                     \
    } while (0)

// Cache of device blocks of one context.  A Davidson solve allocates the
// same buffer sizes every time; a block released to the pool is kept
// instead of being freed (hipFree synchronises the device, ~0.1 ms a call)
// and handed out again for a request of its size (at most 25 % larger).
// Every user of pooled blocks works on the context's one stream, so a block
// handed out again is only touched after the work queued on it before.
struct BlockPool {
    std::multimap<size_t, void *> free_blocks;
    void *get(size_t bytes, size_t *cap) {
        auto it = free_blocks.lower_bound(bytes);
        if (it != free_blocks.end() && it->first <= bytes + bytes / 4) {
            void *q = it->second;
            *cap = it->first;
            free_blocks.erase(it);
            return q;
        }
        void *q = nullptr;
        if (hipMalloc(&q, bytes) != hipSuccess) {
            (void)hipGetLastError();
            clear();  // give the cached blocks back and try once more
            if (hipMalloc(&q, bytes) != hipSuccess)
                throw Error(PCI_E_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
        }
        *cap = bytes;
        return q;
    }
    void put(void *q, size_t cap) { free_blocks.emplace(cap, q); }
    void clear() {
        for (auto &kv : free_blocks) (void)hipFree(kv.second);
        free_blocks.clear();
    }
    ~BlockPool() { clear(); }
};

// Device buffer owned by a context (freed in ~Ctx or on re-allocation), or,
// with pool set, a block of that pool returned to it on release.
template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    size_t cap = 0;            // bytes of the block (pooled buffers)
    BlockPool *pool = nullptr;
    void alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        if (pool) {
            p = (T *)pool->get(count * sizeof(T), &cap);
            n = count;
            return;
        }
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            throw Error(PCI_E_NOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
        }
        n = count;
    }
    void release() {
        if (p) {
            if (pool) pool->put(p, cap);
            else (void)hipFree(p);
        }
        p = nullptr;
        n = 0;
        cap = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
    ~DBuf() { release(); }
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
};

template <typename T>
void upload(DBuf<T> &b, const T *h, size_t count, hipStream_t s) {
    b.alloc(count);
    if (count) PCI_HIP(hipMemcpyAsync(b.p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
}

// Resident Hamiltonian: full symmetric CSR (both triangles) with unshifted
// FP64 values, the per-row length of its lower part, and the FP32 copy with
// the diagonal shifted by -Hmin used by the fast SpMM.
// Buffers of the FP32 dense eigensolver (eigh.hip) between its two phases:
// eigh32_reduce tridiagonalises (the cooperative kernel, which needs the
// whole GPU) or, where that kernel cannot run, solves with ssyevd;
// eigh32_finish computes the eigenvectors of the tridiagonal matrix and
// applies Q (library calls that may share the GPU).
struct Eigh32Work {
    bool coop = false;
    DBuf<float> V, e, tau, y, colpub;
    DBuf<unsigned> flags;
    DBuf<int> abort_flag, info;
};

// Sliced copy of H for the LDS-panel SpMM (k_spmm_sell), one per value
// precision, built on the first SpMM that tries it.  Rows go in chunks of 16
// (one wavefront, 4 lanes per row), columns in panels of P; the elements of
// a chunk's rows inside one panel form a segment, interleaved lane by lane
// and zero-padded to the longest run in the chunk, with 16-bit panel-local
// column indices.  off[c * npan + p] is segment (c, p)'s first pair step
// (128 slots: 2 per lane).
struct SellCopy {
    int P = 0, npan = 0;
    int64_t nchunk = 0, pairs = 0;
    bool built = false, eligible = true;
    float build_ms = 0.f;
    DBuf<int64_t> off;    // nchunk * npan + 1
    DBuf<uint32_t> col;   // pairs * 64 words, two 16-bit columns each
    DBuf<char> val;       // pairs * 128 values (float or double)
    void release() {
        off.release();
        col.release();
        val.release();
        built = false;
        eligible = true;
        pairs = 0;
    }
};

// The determinant space's residual lists, as a matrix built on that space
// sees them (device pointers owned by the context; see Ctx).
struct ResView {
    const int64_t *list_ptr = nullptr;
    const int *list_det = nullptr;
    const int *det_res = nullptr;
    const uint64_t *occ = nullptr;
    int64_t nlists = 0;
    int nres = 0, rsize = 0, ne = 0, W = 0;
};

// Residual-block copy of H, one per value precision.  Two determinants
// couple only if they share Ne-2 positions, so every stored element (i, k)
// belongs to the residual list of the Ne-2 lowest positions they share (the
// pair's canonical list, the one the H build takes it from), and all pairs
// of one list are coupled: H is the sum over lists r of dense blocks B_r on
// the list's determinants, each pair in one block, zero elsewhere.  Block r
// (n = list size) is stored in row chunks of 16 and column steps of 8, in
// the operand order of a 16 x 16 x 4 matrix-core product: value (p, q) of the
// block sits at pair ((p / 16) * steps + q / 8) * 64 + 16 (q mod 4) + p mod
// 16, half (q mod 8) / 4, steps = ceil(n / 8); rows padded to 16, columns
// to 8, with zeros.  One 512-byte wavefront load of a 16 x 8 tile gives lane
// l the values (l mod 16, l / 16) and (l mod 16, 4 + l / 16): the A operands
// of the tile's two 4-column products.  No column indices are stored.  A
// unit of work is (list, range of row chunks).
constexpr int RB_R = 16;  // rows per chunk (one wavefront, 4 lanes per row)
__host__ __device__ inline int64_t rb_index(int p, int q, int steps) {
    return 2 * (((int64_t)(p / RB_R) * steps + (q >> 3)) * 64 + 16 * (q & 3) + (p % RB_R)) + ((q & 7) >> 2);
}
struct RblkCopy {
    bool built = false, eligible = true;
    int nunit = 0, maxn = 0;
    int64_t nvals = 0;
    DBuf<int64_t> boff;   // nlists + 1: first value of each block
    DBuf<int> unit;       // 3 per unit: list, first chunk, end chunk (largest lists first)
    DBuf<char> val;       // nvals values (float or double)
    void release() {
        boff.release();
        unit.release();
        val.release();
        built = false;
        eligible = true;
        nvals = 0;
    }
};

struct Matrix {
    int nd = 0;
    int64_t nnz_full = 0;   // stored elements, both triangles
    int64_t numh = 0;       // stored lower-triangle elements (diag incl.)
    int64_t numzero = 0;
    int64_t ncand = 0;      // determinant pairs evaluated by the build (both triangles)
    double build_ms = 0.0;  // wall time of the build
    double hmin = 0.0;
    DBuf<int64_t> rp;       // nd+1
    DBuf<int> col;          // 0-based columns
    DBuf<double> val;       // FP64, unshifted
    DBuf<int64_t> lowlen;   // per row: entries with col <= row
    DBuf<float> val32;      // FP32, diagonal shifted by -hmin (built on demand)
    float shift32 = 0.f;    // -hmin when some diagonal element is absent (exact zero), else 0
    bool has32 = false;
    bool valid = false;
    bool csf = false;       // H in the CSF basis (rows are CSFs, not determinants)
    // SpMM kernel chosen by the autotuner per (precision 32/64, block width,
    // vector-capable call); -1 = not tuned yet.  Cleared with the matrix.
    signed char spmm_choice[2][65][2];
    // column panels for the cache-blocked SpMM: pan[q] holds, for K = 2 << q
    // panels, the (K+1) x nd element offsets where each row's panel p starts;
    // part is the per-panel partial-sum scratch
    DBuf<int64_t> pan[3];
    DBuf<char> part;
    SellCopy sell[2];   // [0] FP32 values (val32), [1] FP64 values (val); LDS panels
    SellCopy sellg[2];  // the same with 65536-column panels (X gathered from L2)
    // residual-block copies: the space's lists (set by the H build; unset for
    // matrices without one), each determinant's position in its lists, and
    // the list-slot partial sums of one application
    ResView res;
    bool has_res = false;
    DBuf<int> det_pos;   // nd * nres
    DBuf<int> slot_idx;  // nd * nres: list slot of each determinant in each of its lists
    RblkCopy rblk[2];
    DBuf<char> slot;
    // H built straight into the residual blocks (rblk[0] FP32 with the
    // shifted diagonal, rblk[1] FP64): no CSR, so only the block kernels run
    bool blocks_only = false;
    // head-to-tail coupling for a head of hc_nd0 rows: rows hc_nd0..nd-1,
    // their elements with a column < hc_nd0, columns ascending (a compact
    // CSR, rebuilt when a solve uses another head size)
    int hc_nd0 = -1;
    DBuf<int64_t> hc_rp;
    DBuf<int> hc_col;
    DBuf<double> hc_v64;
    DBuf<float> hc_v32;
    // set only during an allocation-free SpMM call: launches take their
    // partial sums from ext (ext_cap bytes) and build no copy
    bool noalloc = false;
    char *ext = nullptr;
    size_t ext_cap = 0;
    void reset_tuning() {
        blocks_only = false;
        hc_nd0 = -1;
        hc_rp.release();
        hc_col.release();
        hc_v64.release();
        hc_v32.release();
        memset(spmm_choice, -1, sizeof spmm_choice);
        for (auto &b : pan) b.release();
        part.release();
        for (auto &b : sell) b.release();
        for (auto &b : sellg) b.release();
        for (auto &b : rblk) b.release();
        det_pos.release();
        slot_idx.release();
        slot.release();
    }
    Matrix() { reset_tuning(); }
};

// Rendezvous of the blocks of one batched Davidson solve.  The cooperative
// head eigensolve spins on a grid barrier and needs every CU, so it may not
// share the GPU with another block's kernels: each block stops there with its
// stream idle (arrive), the blocks then take the GPU one at a time (excl), and
// all continue once every block is through (finish).  A block that fails
// arrives and finishes without waiting, so the others never wait for it.
struct Ctx;
struct BatchGate {
    std::mutex m, excl;
    std::condition_variable cv;
    int n = 0, arrived = 0, finished = 0;
    void arrive(Ctx &c, bool wait);
    void finish(Ctx &c, bool wait);
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // space
    int ns = 0, nso = 0, npos = 0, ne = 0, nd = 0, nc = 0, mj = 0, W = 0, nsu = 0;
    double gj = 0.0;
    bool have_space = false, have_integrals = false;
    DBuf<int> ll, jj, kk, jz, nh;          // 1-based (index 0 unused)
    DBuf<uint64_t> occ;                    // W x nd, word-major (SoA)
    DBuf<int> det_conf;                    // nd
    DBuf<int> det_rank;                    // nd: excitation rank of the configuration vs the first (0..3)
    DBuf<int> det_bucket;                  // nd: 64 * configuration / Nc
    DBuf<int> conf_start;                  // nc+1
    // residual index: determinants differing in at most two positions share a
    // subset of Ne-2 positions (a "residual"); list r holds, ascending, the
    // determinants containing residual r
    int nres = 0, rsize = 0;               // residuals per determinant C(Ne,2), residual size Ne-2
    int64_t nlists = 0;
    DBuf<int64_t> list_ptr;                // nlists+1
    DBuf<int> list_det;                    // nd*nres
    DBuf<uint64_t> list_key;               // packed positions (10 bits each, 0-based)
    DBuf<int> det_res;                     // nd*nres list ids
    std::vector<int> h_conf_start;
    std::vector<int> h_conf_shells;        // nc*ne: each configuration's shells (ascending, with
                                           // repeats) from its first determinant; -1 when it has none
    // CSF basis (kCSF = 1): configuration classes and the resident J = Jm
    // eigenvectors of each class's J^2 block (jbasis's ccj_module)
    struct CsfBasis {
        bool valid = false;
        int nneq = 0, ncsf = 0;
        int64_t nccj = 0;
        std::vector<int> cls;        // per configuration: 0-based class
        std::vector<int> ndcs;       // per class: CSFs
        std::vector<int> ndc_neq;    // per class: determinants of its configurations
        std::vector<int64_t> place;  // per class: first coefficient in ccj
        std::vector<int> mdcs;       // per configuration: first CSF (0-based)
        DBuf<double> ccj;            // per class: ndc_neq x ndcs, column-major
    } csf;
    // integrals
    int nx = 0, kbrt = 0, nlx = 0, gmax = 0, nval = 0, npair = 0;
    DBuf<double> hint, rk, rk2, gaunt;
    // MBPT one-electron effective integrals (Ksig = 1)
    int ksig = 0, kdsig = 0, nmaxs = 0, lmaxs = 0;
    DBuf<double> sg, eps, dg;
    // Ksig = 2: Gint's screening coefficients and the two-electron sigma table
    bool have_sigma2 = false;
    int ksym = 1, kmaxs = 0, nsum = 0;
    double scr[10] = {};
    int64_t s2mask = 0;
    DBuf<int64_t> s2key;
    DBuf<double> s2v;
    DBuf<int64_t> quad;
    DBuf<int> err;
    // learned correction network
    struct MlpNet *mlp = nullptr;
    // matrices
    Matrix H;
    Matrix J;    // J^2, lower triangle (pconf's Jsq CSR)
    Matrix Jf;   // J^2 with both triangles, for block products (built on demand)
    // J selection (Kv = 3): target J and the J terms of the space
    double xj_av = 0.0;
    std::vector<int> jt;
    // scratch
    DBuf<char> work;
    double *pin = nullptr;  // pinned host staging for the small per-iteration transfers
    void *blas = nullptr;  // rocBLAS handle on the context stream, created on first use
    BlockPool pool;        // blocks of the Davidson and eigensolver buffers, reused across solves
    int cu_count = 0;      // device properties the eigensolver needs (read once)
    bool coop_launch = false;
    int trd_fallbacks = 0;  // head eigensolves finished by ssyevd after a k_trd barrier ran out
    size_t lds_per_block = 0;
    struct BatchGate *gate = nullptr;  // set while the context is one block of a batched solve
    bool gate_arrived = false, gate_finished = false;
    size_t pin_n = 0;
    Tables tables() const;
    ~Ctx();
};

void mlp_free(Ctx &c);
// formh.hip
void build_matrix(Ctx &c, bool j2, Matrix &M);
void ensure_fp32(Ctx &c);
void scan_exclusive(const int64_t *in, int64_t *out, int64_t n, hipStream_t s, DBuf<char> &work);
// spmm.hip
// padded: every X row may be read up to the next 16-byte boundary past nvec
// The first call for a (precision, nvec) on a matrix times the candidate
// kernels on that call's operands and keeps the fastest (autotuner).
void spmm_f64(Ctx &c, const double *X, int ldx, double *Y, int ldy, int nvec, hipStream_t s, bool padded);
void spmm_f32(Ctx &c, const float *X, int ldx, float *Y, int ldy, int nvec, hipStream_t s, bool padded);
// Head-to-tail coupling: Y[i, :] = sum_{k < nd0} H_ik X[k, :] for rows
// i >= nd0 (X rows 16-byte aligned and padded), using the FP64 or the FP32
// matrix.  head_coupling_setup builds the compact rows for this nd0 first.
void head_coupling_setup(Ctx &c, int nd0, hipStream_t s);
void head_couple_f64(const Ctx &c, int nd0, const double *X, int ldx, double *Y, int ldy, int nvec, hipStream_t s);
void head_couple_f32(const Ctx &c, int nd0, const float *X, int ldx, float *Y, int ldy, int nvec, hipStream_t s);
// Init4's dense head block Z (nd0 x nd0, column-major, zeroed by the caller)
// of H - Hmin I, and the shifted diagonal D (nd), from the resident H in
// whichever storage it has.
void head_block_and_diag(const Ctx &c, int nd0, double *Z, double *D, hipStream_t s);
// residual-block layout of M's lists (block offsets, units largest first,
// nvals, maxn; uploads boff and unit) and the determinants' list positions
// (det_pos, slot_idx), shared by the copy built from the CSR and the direct
// block build
int64_t rblk_layout(Matrix &M, RblkCopy &B, hipStream_t s);
void rblk_positions(Matrix &M, hipStream_t s);
// rows [row0, row0 + nrows) of the lower triangle of a blocks-only H, as
// copy_lower_rows gives them from the CSR
void rblk_lower_rows(Ctx &c, Matrix &M, int row0, int nrows, int64_t *rowptr, int32_t *col, double *val,
                     int64_t cap);
// Y = (M + shift I) X for any resident matrix M (FP64), e.g. J^2
void spmm_matrix_f64(Matrix &M, const double *X, int ldx, double *Y, int ldy, int nvec, double shift, hipStream_t s,
                     bool padded);
// full-storage copy of the lower-triangle J^2 (c.Jf)
void ensure_jfull(Ctx &c);
// name of the kernel variant in use for (precision, nvec), or "" if untuned
const char *spmm_variant_name(const Ctx &c, int prec, int nvec);
// this device's copy of the allocation-free SpMM's scratch array (once per device)
void scratch_setup(int device);
double spmm_bytes(const Ctx &c, int nvec, int prec);

}  // namespace pci

struct pci_ctx {
    pci::Ctx c;
};
