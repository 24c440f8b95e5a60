// Context lifetime, space and integral upload, error reporting.
//
// set_space replaces the InitFormH broadcast (src/pconf.F90:1049-1158) and
// FormBarr (src/determinants.f90:839-864): determinants are bit-packed into
// 64-bit words, stored word-major (word w of determinant n at w*Nd + n) so a
// wavefront reading 64 consecutive determinants issues coalesced loads.
// It also builds the residual index that replaces the per-row CompD screen of
// FormH (src/pconf.F90:1236-1247): two determinants differ in at most two
// spin-orbitals exactly when they share Ne-2 positions, so the determinants
// a row couples to are found in the C(Ne,2) inverted lists of its own
// (Ne-2)-position subsets instead of by scanning whole configuration blocks.
//
// set_integrals replaces BuildHintLUT/BuildGintHash/BuildGauntLUT
// (src/integrals.f90:604-715).  The open-addressing hash of (iac, ibd) keys
// becomes a dense table indexed by the canonical shell quadruple
// (pair(na,nc), pair(nb,nd)) holding the base of a contiguous run of R^k
// indexed by k, so a lookup is two dependent loads with no probing.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

#include <rocblas/rocblas.h>

#include "internal.h"

namespace pci {

static thread_local std::string g_last_error;
void set_last_error(const std::string &m) { g_last_error = m; }

Tables Ctx::tables() const {
    Tables T;
    T.jz = jz.p; T.nh = nh.p; T.ll = ll.p; T.jj = jj.p; T.kk = kk.p;
    T.hint = hint.p; T.quad = quad.p; T.rk = rk.p; T.rk2 = kbrt ? rk2.p : rk.p; T.gaunt = gaunt.p;
    T.nso = nso; T.nx = nx; T.nval = nval; T.npair = npair;
    T.kbrt = kbrt; T.nlx = nlx; T.gmax = gmax; T.ne = ne; T.mj = mj; T.gj = gj;
    T.err = err.p;
    T.ksig = ksig; T.kdsig = kdsig; T.nmaxs = nmaxs; T.lmaxs = lmaxs;
    T.sg = sg.p; T.eps = eps.p; T.dg = dg.p;
    T.ksym = ksym; T.kmaxs = kmaxs; T.nsum = nsum;
    for (int k = 0; k < 10; k++) T.scr[k] = scr[k];
    T.s2mask = s2mask; T.s2key = s2key.p; T.s2v = s2v.p;
    return T;
}

void mlp_free(Ctx &c);
Ctx::~Ctx() {
    mlp_free(*this);
    if (blas) (void)rocblas_destroy_handle((rocblas_handle)blas);
    if (pin) (void)hipHostFree(pin);
    if (stream) (void)hipStreamDestroy(stream);
}

// ---------------------------------------------------------------------------
// residual index (host): every determinant contributes its C(Ne, Ne-2)
// subsets of Ne-2 positions; (key, det, r) triples are sorted so each list
// holds ascending determinants.
// ---------------------------------------------------------------------------
static void build_residual_index(Ctx &c, const int32_t *iarr) {
    const int ne = c.ne, nd = c.nd;
    const int R = ne > 2 ? ne - 2 : 0;
    PCI_REQUIRE(R <= 6, PCI_E_UNSUPPORTED, "more than 8 valence electrons");
    std::vector<std::vector<int>> combos;
    std::vector<int> cur;
    // all R-subsets of {0..ne-1} in lexicographic order
    std::vector<int> idx(R);
    for (int i = 0; i < R; i++) idx[i] = i;
    while (true) {
        combos.push_back(idx);
        int i = R - 1;
        while (i >= 0 && idx[i] == ne - R + i) i--;
        if (i < 0) break;
        idx[i]++;
        for (int j = i + 1; j < R; j++) idx[j] = idx[j - 1] + 1;
    }
    if (R == 0) combos.assign(1, std::vector<int>());
    const int nres = (int)combos.size();
    struct Ent { uint64_t key; int det; int r; };
    std::vector<Ent> e((size_t)nd * nres);
    for (int64_t n = 0; n < nd; n++)
        for (int r = 0; r < nres; r++) {
            uint64_t key = 0;
            for (int q = 0; q < R; q++) key |= (uint64_t)(iarr[n * ne + combos[r][q]] - 1) << (10 * q);
            e[(size_t)n * nres + r] = {key, (int)n, r};
        }
    std::sort(e.begin(), e.end(), [](const Ent &a, const Ent &b) {
        return a.key < b.key || (a.key == b.key && a.det < b.det);
    });
    std::vector<int64_t> ptr;
    std::vector<uint64_t> keys;
    std::vector<int> dets(e.size()), dres((size_t)nd * nres);
    for (size_t i = 0; i < e.size(); i++) {
        if (i == 0 || e[i].key != e[i - 1].key) {
            ptr.push_back((int64_t)i);
            keys.push_back(e[i].key);
        }
        dets[i] = e[i].det;
        dres[(size_t)e[i].det * nres + e[i].r] = (int)(keys.size() - 1);
    }
    ptr.push_back((int64_t)e.size());
    c.nres = nres;
    c.rsize = R;
    c.nlists = (int64_t)keys.size();
    hipStream_t s = c.stream;
    upload(c.list_ptr, ptr.data(), ptr.size(), s);
    upload(c.list_key, keys.data(), keys.size(), s);
    upload(c.list_det, dets.data(), dets.size(), s);
    upload(c.det_res, dres.data(), dres.size(), s);
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

extern "C" const char *pci_gpu_last_error(void) { return g_last_error.c_str(); }

extern "C" int pci_gpu_create(int device, pci_ctx **out) {
    API_BEGIN
    PCI_REQUIRE(out != nullptr, PCI_E_ARG, "pci_gpu_create: null output pointer");
    int n = 0;
    PCI_HIP(hipGetDeviceCount(&n));
    PCI_REQUIRE(device >= 0 && device < n, PCI_E_ARG, "pci_gpu_create: no such device");
    PCI_HIP(hipSetDevice(device));
    scratch_setup(device);  // the allocation-free SpMM's scratch on this device
    pci_ctx *p = new pci_ctx();
    p->c.device = device;
    hipError_t e = hipStreamCreateWithFlags(&p->c.stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete p;
        throw Error(PCI_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    p->c.err.alloc(4);
    PCI_HIP(hipMemsetAsync(p->c.err.p, 0, 4 * sizeof(int), p->c.stream));
    *out = p;
    API_END
}

extern "C" int pci_gpu_destroy(pci_ctx *ctx) {
    API_BEGIN
    if (ctx) {
        (void)hipSetDevice(ctx->c.device);
        (void)hipStreamSynchronize(ctx->c.stream);
        delete ctx;
    }
    API_END
}

extern "C" int pci_gpu_set_space(pci_ctx *ctx, int32_t ns, int32_t nso, const int32_t *ll,
                                 const int32_t *jj, const int32_t *kk, int32_t npos,
                                 const int32_t *jz, const int32_t *nh, int32_t ne, int32_t nd,
                                 const int32_t *iarr, int32_t nc, const int32_t *ndc, int32_t mj,
                                 double gj) {
    API_BEGIN
    PCI_REQUIRE(ctx, PCI_E_ARG, "null context");
    Ctx &c = ctx->c;
    PCI_HIP(hipSetDevice(c.device));
    PCI_REQUIRE(ns > 0 && nso >= 0 && nso < ns && npos > 0 && ne > 0 && ne <= 16 && nd > 0 && nc > 0,
                PCI_E_ARG, "pci_gpu_set_space: bad dimensions (1 <= Ne <= 16 required)");
    PCI_REQUIRE(ll && jj && kk && jz && nh && iarr && ndc, PCI_E_ARG, "pci_gpu_set_space: null array");
    c.pool.clear();  // blocks sized for the previous space
    int64_t sum = 0;
    for (int i = 0; i < nc; i++) {
        PCI_REQUIRE(ndc[i] >= 0, PCI_E_ARG, "negative Ndc");
        sum += ndc[i];
    }
    PCI_REQUIRE(sum == nd, PCI_E_ARG, "sum(Ndc) != Nd");
    int maxpos = 0;
    for (int64_t i = 0; i < (int64_t)nd * ne; i++) {
        PCI_REQUIRE(iarr[i] >= 1 && iarr[i] <= npos, PCI_E_ARG, "Iarr position out of range");
        maxpos = std::max(maxpos, (int)iarr[i]);
    }
    for (int i = 0; i < npos; i++)
        PCI_REQUIRE(nh[i] >= 1 && nh[i] <= ns, PCI_E_ARG, "Nh shell out of range");
    const int W = (maxpos + 63) / 64;
    PCI_REQUIRE(W <= 16, PCI_E_UNSUPPORTED, "more than 1024 spin-orbital positions in use");
    c.ns = ns; c.nso = nso; c.npos = npos; c.ne = ne; c.nd = nd; c.nc = nc; c.mj = mj; c.gj = gj; c.W = W;
    hipStream_t s = c.stream;
    // 1-based tables
    std::vector<int> t(ns + 1, 0);
    for (int i = 0; i < ns; i++) t[i + 1] = ll[i];
    upload(c.ll, t.data(), ns + 1, s);
    for (int i = 0; i < ns; i++) t[i + 1] = jj[i];
    upload(c.jj, t.data(), ns + 1, s);
    for (int i = 0; i < ns; i++) t[i + 1] = kk[i];
    upload(c.kk, t.data(), ns + 1, s);
    std::vector<int> p(npos + 1, 0);
    for (int i = 0; i < npos; i++) p[i + 1] = jz[i];
    upload(c.jz, p.data(), npos + 1, s);
    for (int i = 0; i < npos; i++) p[i + 1] = nh[i];
    upload(c.nh, p.data(), npos + 1, s);
    c.nsu = 0;
    for (int64_t i = 0; i < (int64_t)nd * ne; i++) c.nsu = std::max(c.nsu, (int)nh[iarr[i] - 1]);
    // bit-packed determinants, word-major
    std::vector<uint64_t> occ((size_t)W * nd, 0);
    for (int64_t n = 0; n < nd; n++)
        for (int q = 0; q < ne; q++) {
            const int pos0 = iarr[n * ne + q] - 1;
            occ[(size_t)(pos0 >> 6) * nd + n] |= 1ull << (pos0 & 63);
        }
    for (int64_t n = 0; n < nd; n++) {
        int pc = 0;
        for (int w = 0; w < W; w++) pc += __builtin_popcountll(occ[(size_t)w * nd + n]);
        PCI_REQUIRE(pc == ne, PCI_E_ARG, "determinant with a repeated position");
    }
    upload(c.occ, occ.data(), occ.size(), s);
    // configuration blocks
    c.h_conf_start.assign(nc + 1, 0);
    std::vector<int> dc(nd);
    for (int i = 0; i < nc; i++) {
        c.h_conf_start[i + 1] = c.h_conf_start[i] + ndc[i];
        for (int n = c.h_conf_start[i]; n < c.h_conf_start[i + 1]; n++) dc[n] = i;
    }
    upload(c.conf_start, c.h_conf_start.data(), nc + 1, s);
    upload(c.det_conf, dc.data(), nd, s);
    // excitation rank of each configuration against the first one (shell-level
    // CompD, src/determinants.f90:748-779) and its position bucket: inputs of
    // the learned correction
    {
        std::vector<int> cs0(ne), csi(ne), rank_c(nc, 3), rk(nd), bk(nd);
        int first = -1;
        for (int i = 0; i < nc && first < 0; i++)
            if (ndc[i] > 0) first = i;
        for (int q = 0; q < ne; q++) cs0[q] = nh[iarr[(int64_t)c.h_conf_start[first] * ne + q] - 1];
        for (int i = 0; i < nc; i++) {
            if (ndc[i] == 0) continue;
            for (int q = 0; q < ne; q++) csi[q] = nh[iarr[(int64_t)c.h_conf_start[i] * ne + q] - 1];
            int a = 0, b = 0, ni = 0, nj = 0;
            while (a < ne && b < ne) {
                if (cs0[a] == csi[b]) { a++; b++; }
                else if (cs0[a] > csi[b]) { nj++; b++; }
                else { ni++; a++; }
            }
            ni += ne - a;
            rank_c[i] = std::min(3, std::max(ni, nj + (ne - b)));
        }
        for (int n = 0; n < nd; n++) {
            rk[n] = rank_c[dc[n]];
            bk[n] = (int)((int64_t)64 * dc[n] / nc);
        }
        upload(c.det_rank, rk.data(), nd, s);
        upload(c.det_bucket, bk.data(), nd, s);
    }
    // each configuration's shells from its first determinant (positions are
    // ascending, so are their shells): the CSF build pairs configurations by
    // them
    c.h_conf_shells.assign((size_t)nc * ne, -1);
    for (int i = 0; i < nc; i++)
        if (ndc[i] > 0)
            for (int q = 0; q < ne; q++)
                c.h_conf_shells[(size_t)i * ne + q] = nh[iarr[(int64_t)c.h_conf_start[i] * ne + q] - 1];
    c.csf.valid = false;
    c.csf.ccj.release();
    PCI_REQUIRE(maxpos <= 1024, PCI_E_UNSUPPORTED, "more than 1024 spin-orbital positions");
    build_residual_index(c, iarr);
    c.have_space = true;
    c.H.valid = false;
    c.J.valid = false;
    c.Jf.valid = false;
    API_END
}

extern "C" int pci_gpu_set_integrals(pci_ctx *ctx, int32_t nx, int32_t nhint, const double *rint1,
                                     const int32_t *iint1, int64_t ngint, int32_t kbrt,
                                     const double *rint2, const int32_t *iint2, const int32_t *iint3,
                                     int32_t nlx, int32_t ngaunt, const int32_t *in, const double *gnt) {
    API_BEGIN
    PCI_REQUIRE(ctx, PCI_E_ARG, "null context");
    Ctx &c = ctx->c;
    PCI_REQUIRE(c.have_space, PCI_E_ARG, "pci_gpu_set_integrals: call pci_gpu_set_space first");
    PCI_HIP(hipSetDevice(c.device));
    PCI_REQUIRE(nx > 0 && nx <= 46340 && nhint >= 0 && ngint >= 0 && ngaunt > 0 && nlx >= 0, PCI_E_ARG,
                "pci_gpu_set_integrals: bad dimensions");
    PCI_REQUIRE(c.ns - c.nso <= nx, PCI_E_ARG, "more valence shells than Nx");
    hipStream_t s = c.stream;
    c.nx = nx; c.kbrt = kbrt; c.nlx = nlx;
    // one-electron LUT
    std::vector<double> h((size_t)nx * nx + 1, 0.0);
    for (int i = 0; i < nhint; i++) {
        PCI_REQUIRE(iint1[i] >= 1 && iint1[i] <= nx * nx, PCI_E_ARG, "Iint1 index out of range");
        h[iint1[i]] = rint1[i];
    }
    upload(c.hint, h.data(), h.size(), s);
    // two-electron quad table
    const int nval = c.ns - c.nso;
    c.nval = nval;
    c.npair = nval * (nval + 1) / 2;
    const int64_t nq = (int64_t)c.npair * c.npair;
    std::vector<int64_t> qb(nq, -1);
    std::vector<int> qk(nq, -1);  // kmax of the run
    std::vector<int> jjh(c.ns + 1);
    PCI_HIP(hipMemcpyAsync(jjh.data(), c.jj.p, (c.ns + 1) * sizeof(int), hipMemcpyDeviceToHost, s));
    PCI_HIP(hipStreamSynchronize(s));
    auto pair = [nval](int a, int b) -> int64_t { return (int64_t)a * nval - (int64_t)a * (a - 1) / 2 + (b - a); };
    const int64_t nx2 = (int64_t)nx * nx;
    std::vector<int64_t> keyq(ngint, -1);
    std::vector<int> keyk(ngint, 0);
    for (int64_t i = 0; i < ngint; i++) {
        const int64_t iac = (int64_t)iint2[i] - 1, ibd = (int64_t)iint3[i] - 1;
        if (iac < 0 || ibd < 0) continue;
        const int k = (int)(iac / nx2);
        const int a = (int)((iac / nx) % nx), cc = (int)(iac % nx);
        const int b = (int)(ibd / nx), d = (int)(ibd % nx);
        if (b >= nx || a >= nval || cc >= nval || b >= nval || d >= nval) continue;
        // keys Gint can generate: na <= nc, nb <= nd, na <= nb, (na == nb -> nc <= nd)
        if (a > cc || b > d || a > b || (a == b && cc > d)) continue;
        const int ja = jjh[a + c.nso + 1], jc = jjh[cc + c.nso + 1], jb = jjh[b + c.nso + 1], jd = jjh[d + c.nso + 1];
        const int kmax = std::min((ja + jc) / 2, (jb + jd) / 2);
        if (k > kmax) continue;
        const int64_t q = pair(a, cc) * c.npair + pair(b, d);
        keyq[i] = q;
        keyk[i] = k;
        qk[q] = kmax;
    }
    int64_t tot = 0;
    for (int64_t q = 0; q < nq; q++)
        if (qk[q] >= 0) {
            qb[q] = tot;
            tot += qk[q] + 1;
        }
    const double qnan = std::numeric_limits<double>::quiet_NaN();
    std::vector<double> r1(std::max<int64_t>(tot, 1), qnan), r2;
    if (kbrt) r2.assign(std::max<int64_t>(tot, 1), qnan);
    const int w = kbrt ? 2 : 1;
    for (int64_t i = 0; i < ngint; i++) {
        if (keyq[i] < 0) continue;
        const int64_t slot = qb[keyq[i]] + keyk[i];
        if (r1[slot] == r1[slot]) continue;  // first occurrence wins, as in FindGint
        r1[slot] = rint2[w * i];
        if (kbrt) r2[slot] = rint2[w * i + 1];
    }
    upload(c.quad, qb.data(), qb.size(), s);
    upload(c.rk, r1.data(), r1.size(), s);
    if (kbrt) upload(c.rk2, r2.data(), r2.size(), s);
    // Gaunt LUT
    int mx = 0;
    for (int i = 0; i < ngaunt; i++) {
        PCI_REQUIRE(in[i] >= 0, PCI_E_ARG, "negative Gaunt index");
        mx = std::max(mx, (int)in[i]);
    }
    std::vector<double> g((size_t)mx + 1, 0.0);
    for (int i = 0; i < ngaunt; i++) g[in[i]] = gnt[i];
    c.gmax = mx;
    upload(c.gaunt, g.data(), g.size(), s);
    PCI_HIP(hipMemsetAsync(c.err.p, 0, 4 * sizeof(int), s));
    PCI_HIP(hipStreamSynchronize(s));
    c.have_integrals = true;
    c.ksig = 0;  // new integrals: effective integrals are set again if wanted
    c.H.valid = false;
    c.J.valid = false;
    c.Jf.valid = false;
    API_END
}

// RintS (SGC.CON part, src/pconf.F90:715-762) + BuildHintSLUT
// (src/integrals.f90:676-687) as device tables: (Rsig, Dsig, Esig) at each
// Hint index, NaN where SGC.CON has no entry; Eps by shell; FormD's Diag.
extern "C" int pci_gpu_set_sigma(pci_ctx *ctx, int32_t ksig, int32_t kdsig, int32_t nmaxs, int32_t lmaxs,
                                 int32_t nhints, const int32_t *iint1s, const double *rsig, const double *dsig,
                                 const double *esig, const double *eps, const double *diag) {
    API_BEGIN
    PCI_REQUIRE(ctx, PCI_E_ARG, "null context");
    Ctx &c = ctx->c;
    PCI_REQUIRE(c.have_space && c.have_integrals, PCI_E_ARG,
                "pci_gpu_set_sigma: call pci_gpu_set_space and pci_gpu_set_integrals first");
    PCI_REQUIRE(ksig >= 0 && ksig <= 2, PCI_E_ARG, "pci_gpu_set_sigma: Ksig must be 0, 1 or 2");
    PCI_HIP(hipSetDevice(c.device));
    hipStream_t s = c.stream;
    c.H.valid = false;
    c.Jf.valid = false;
    c.have_sigma2 = false;
    if (ksig == 0) {
        c.ksig = 0;
        c.kdsig = 0;
        return PCI_OK;
    }
    PCI_REQUIRE(nhints >= 0 && nmaxs >= 0 && nmaxs <= c.ns && lmaxs >= 0, PCI_E_ARG,
                "pci_gpu_set_sigma: bad dimensions");
    PCI_REQUIRE(nhints == 0 || (iint1s && rsig && dsig && esig), PCI_E_ARG, "pci_gpu_set_sigma: null table");
    PCI_REQUIRE(eps, PCI_E_ARG, "pci_gpu_set_sigma: null Eps");
    PCI_REQUIRE(kdsig == 0 || diag, PCI_E_ARG, "pci_gpu_set_sigma: Kdsig needs FormD's Diag");
    const int64_t nx2 = (int64_t)c.nx * c.nx;
    std::vector<double> g(3 * (size_t)(nx2 + 1), 0.0);
    for (int64_t i = 0; i <= nx2; i++) g[3 * i] = std::numeric_limits<double>::quiet_NaN();
    for (int i = 0; i < nhints; i++) {
        PCI_REQUIRE(iint1s[i] >= 1 && iint1s[i] <= nx2, PCI_E_ARG, "Iint1S index out of range");
        g[3 * (size_t)iint1s[i]] = rsig[i];
        g[3 * (size_t)iint1s[i] + 1] = dsig[i];
        g[3 * (size_t)iint1s[i] + 2] = esig[i];
    }
    upload(c.sg, g.data(), g.size(), s);
    std::vector<double> e(c.ns + 1, 0.0);
    for (int i = 1; i <= c.ns; i++) e[i] = eps[i - 1];
    upload(c.eps, e.data(), e.size(), s);
    if (kdsig != 0) upload(c.dg, diag, (size_t)c.nd, s);
    PCI_HIP(hipStreamSynchronize(s));
    c.ksig = ksig;
    c.kdsig = kdsig;
    c.nmaxs = nmaxs;
    c.lmaxs = lmaxs;
    API_END
}

extern "C" int pci_gpu_set_sigma2(pci_ctx *ctx, int32_t ksym, int32_t kmax, int32_t nsum, const double *scr,
                                  int64_t ngints, const int32_t *iint2s, const int32_t *iint3s, const double *rsig2,
                                  const double *dsig2, const double *esig2) {
    API_BEGIN
    PCI_REQUIRE(ctx, PCI_E_ARG, "null context");
    Ctx &c = ctx->c;
    PCI_REQUIRE(c.ksig == 2, PCI_E_ARG, "pci_gpu_set_sigma2: call pci_gpu_set_sigma with Ksig = 2 first");
    PCI_REQUIRE(scr && kmax >= 0 && ngints >= 0 && (ngints == 0 || (iint2s && iint3s && rsig2 && dsig2 && esig2)),
                PCI_E_ARG, "pci_gpu_set_sigma2: bad arguments");
    PCI_HIP(hipSetDevice(c.device));
    // open-addressing table of the (Iint2S, Iint3S) keys, at most half full
    int64_t cap = 16;
    while (cap < 2 * ngints) cap *= 2;
    std::vector<int64_t> key((size_t)cap, -1);
    std::vector<double> v(3 * (size_t)cap, 0.0);
    for (int64_t i = 0; i < ngints; i++) {
        PCI_REQUIRE(iint2s[i] >= 1 && iint3s[i] >= 1, PCI_E_ARG, "pci_gpu_set_sigma2: bad SCRC.CON index");
        const int64_t k = ((int64_t)iint2s[i] << 32) | (uint32_t)iint3s[i];
        int64_t slot = s2_hash(k, cap - 1);
        while (key[slot] != -1 && key[slot] != k) slot = (slot + 1) & (cap - 1);
        key[slot] = k;  // keys are unique in SCRC.CON
        v[3 * slot] = rsig2[i];
        v[3 * slot + 1] = dsig2[i];
        v[3 * slot + 2] = esig2[i];
    }
    upload(c.s2key, key.data(), key.size(), c.stream);
    upload(c.s2v, v.data(), v.size(), c.stream);
    PCI_HIP(hipStreamSynchronize(c.stream));
    c.ksym = ksym;
    c.kmaxs = kmax;
    c.nsum = nsum;
    for (int k = 0; k < 10; k++) c.scr[k] = scr[k];
    c.s2mask = cap - 1;
    c.have_sigma2 = true;
    c.H.valid = false;
    c.Jf.valid = false;
    API_END
}

extern "C" int pci_gpu_mem_usage(pci_ctx *ctx, int64_t *bytes) {
    API_BEGIN
    PCI_REQUIRE(ctx && bytes, PCI_E_ARG, "null argument");
    const Ctx &c = ctx->c;
    int64_t b = c.ll.bytes() + c.jj.bytes() + c.kk.bytes() + c.jz.bytes() + c.nh.bytes() + c.occ.bytes() +
                c.det_conf.bytes() + c.conf_start.bytes() + c.list_ptr.bytes() + c.list_det.bytes() +
                c.list_key.bytes() + c.det_res.bytes() +
                c.hint.bytes() + c.rk.bytes() + c.rk2.bytes() + c.gaunt.bytes() + c.quad.bytes() +
                c.work.bytes();
    for (const Matrix *m : {&c.H, &c.J, &c.Jf}) {
        b += m->rp.bytes() + m->col.bytes() + m->val.bytes() + m->lowlen.bytes() + m->val32.bytes();
        b += m->part.bytes() + m->det_pos.bytes() + m->slot_idx.bytes() + m->slot.bytes();
        for (const DBuf<int64_t> &q : m->pan) b += q.bytes();
        for (int p = 0; p < 2; p++) {
            for (const SellCopy *S : {&m->sell[p], &m->sellg[p]}) b += S->off.bytes() + S->col.bytes() + S->val.bytes();
            b += m->rblk[p].boff.bytes() + m->rblk[p].unit.bytes() + m->rblk[p].val.bytes();
        }
    }
    *bytes = b;
    API_END
}
