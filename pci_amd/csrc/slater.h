// Slater-Condon matrix elements on bit-packed determinants.
//
// A determinant is W 64-bit words; spin-orbital position p (1-based) is bit
// (p-1) % 64 of word (p-1) / 64.  For a pair (det1, det2) the excitation
// level is popcount(det1 ^ det2) / 2 and the differing positions are the set
// bits of det1 & ~det2 (det1 only) and det2 & ~det1 (det2 only).
//
// This restates, on bitstrings, what the reference computes on sorted
// integer lists:
//   Rspq_phase1/2   src/determinants.f90:595-725  -> excitation, positions, sign
//   Hmltn           src/pconf.F90:1653-1692        -> matrix element
//   Gint            src/integrals.f90:153-293
//   Gaunt           src/integrals.f90:509-561
//   Hint, HintS     src/integrals.f90:122-151, 321-369
//   GintS           src/integrals.f90:371-493 (with Gint's screening, 262-271)
//   F_J2, Plj       src/formj2.f90:13-67
// The Rspq sign (-1)^(|j2-i2| + |j3-i3|) uses the 1-based ranks of the
// differing positions inside each sorted determinant; here a rank is
// popcount of the determinant below that position plus one.  Terms are
// summed in the reference's order with contraction disabled, so every value
// is bit-identical to the reference's FP64 result (and its exact-zero
// decisions are reproduced).
#pragma once
#include <stdint.h>

#ifndef PCI_HD
#if defined(__HIPCC__) || defined(__HIP__)
#define PCI_HD __host__ __device__ inline
#else
#define PCI_HD inline
#endif
#endif

namespace pci {

// Device-resident lookup tables for one space (all pointers are device
// pointers; the same struct is used with host pointers in host tests).
struct Tables {
    const int *jz;          // 2m per position, index 1..npos
    const int *nh;          // shell per position, index 1..npos
    const int *ll;          // l per shell, index 1..ns
    const int *jj;          // 2j per shell
    const int *kk;          // kappa per shell
    const double *hint;     // HintLUT, index Nx*(na-Nso-1)+(nb-Nso), 1..Nx*Nx
    const int64_t *quad;    // base of the k-run of each canonical shell quadruple, -1 absent
    const double *rk;       // Coulomb R^k by (quad base + k); NaN where absent
    const double *rk2;      // Breit part (kbrt != 0)
    const double *gaunt;    // GauntLUT 0..gmax
    int nso, nx, nval;      // core shells, index radix, valence shells (Nsu-Nso)
    int npair;              // nval*(nval+1)/2
    int kbrt, nlx, gmax, ne, mj;
    double gj;
    int *err;               // set to 1 on a missing integral
    // MBPT one-electron effective integrals (Ksig = 1)
    int ksig, kdsig, nmaxs, lmaxs;
    const double *sg;       // (Rsig, Dsig, Esig) per Hint index, Rsig NaN where absent
    const double *eps;      // orbital energies, index 1..ns
    const double *dg;       // FormD's Diag per determinant (0-based), Kdsig != 0
    // MBPT screening of Gint and two-electron sigma (Ksig = 2, SCRC.CON)
    int ksym, kmaxs, nsum;
    double scr[10];         // screening coefficient of multipole k = 0..9
    int64_t s2mask;         // slots - 1 (a power of two less one)
    const int64_t *s2key;   // per slot: Iint2S << 32 | Iint3S, -1 empty
    const double *s2v;      // per slot: (R, D, E)
};

// Slot of an (Iint2S, Iint3S) key in the sigma table: a multiplicative hash
// of the packed key, linear probing (the host builds the table the same way).
PCI_HD int64_t s2_hash(int64_t key, int64_t mask) {
    return (int64_t)(((uint64_t)key * 0x9E3779B97F4A7C15ull) >> 20) & mask;
}

PCI_HD int popc64(uint64_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __popcll(x);
#else
    return __builtin_popcountll(x);
#endif
}

PCI_HD int ctz64(uint64_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __ffsll((unsigned long long)x) - 1;
#else
    return __builtin_ctzll(x);
#endif
}

// Gaunt(k; j1 m1, j2 m2) with all angular arguments doubled (integers).
PCI_HD double gaunt(const Tables &T, int k, int j1, int m1, int j2, int m2) {
    int is = 1;
    int dm = m2 - m1;
    int im = (dm < 0 ? -dm : dm) / 2;
    if (k >= im) {
        if (k == 0) {
            if (j1 == j2) return 1.0;
        } else {
            if (j2 < j1) {
                int x = j2; j2 = j1; j1 = x;
                x = m2; m2 = m1; m1 = x;
                if (im & 1) is = -is;
            }
            if (m1 <= 0) {
                m1 = -m1; m2 = -m2;
                int ij = k + (j1 - j2) / 2;
                if (ij & 1) is = -is;
            }
            const int ib1 = 2 * T.nlx + 1, ib2 = ib1 * ib1;
            const long ind = (long)ib2 * ((long)ib2 * k + (ib1 * j1 + j2)) + ib1 * ((j1 + m1) / 2) + (j2 + m2) / 2;
            if (ind >= 0 && ind <= T.gmax) return T.gaunt[ind] * (double)is;
        }
    }
    return 0.0 * (double)is;
}

// HintS(na, nb, n0) (src/integrals.f90:321-369) with Kexn = 1: the sigma
// correction of the pair na <= nb, extrapolated in E_k + Eps(n0) - Esig when
// Kdsig != 0 and Dsig != 0 (ek: Diag of the column determinant).
PCI_HD double hints(const Tables &T, int na, int nb, int n0, double ek) {
#pragma clang fp contract(off)
    double e = 0.0;
    if (T.ll[na] > T.lmaxs || nb > T.nmaxs) return e;
    const double *g = T.sg + 3 * (T.nx * (na - T.nso - 1) + (nb - T.nso));
    e = g[0];
    if (e != e) {  // "HintS: integral is absent"
        *T.err = 1;
        return 0.0;
    }
    const double d = g[1];
    if ((double)T.kdsig * d != 0.0) {
        const double de = (ek + T.eps[n0] - g[2]);
        const double d1 = d / e;
        double dd = d1 * de;
        if (dd > 0.5) dd = 0.5;
        if (d1 > 0.0) e = e / (1.0 - dd);   // normal variant
        else e = e * (1.0 + dd);            // anomalous: two-side extrapolation
    }
    return e;
}

PCI_HD double hint(const Tables &T, int ia, int ib, double ek = 0.0) {
#pragma clang fp contract(off)
    double e = 0.0;
    if (T.jz[ia] == T.jz[ib]) {
        int na = T.nh[ia], nb = T.nh[ib];
        const int n0 = na;
        if (T.kk[na] == T.kk[nb]) {
            if (na > nb) { int t = na; na = nb; nb = t; }
            e = T.hint[T.nx * (na - T.nso - 1) + (nb - T.nso)];
            if (T.ksig != 0) e = e + hints(T, na, nb, n0, ek);
        }
    }
    return e;
}

PCI_HD int64_t quad_index(const Tables &T, int na, int nc, int nb, int nd) {
    // canonical: na <= nc, nb <= nd (0-based valence shells after -Nso-1)
    int a = na - T.nso - 1, c = nc - T.nso - 1, b = nb - T.nso - 1, d = nd - T.nso - 1;
    int64_t pac = (int64_t)a * T.nval - (int64_t)a * (a - 1) / 2 + (c - a);
    int64_t pbd = (int64_t)b * T.nval - (int64_t)b * (b - 1) / 2 + (d - b);
    return pac * T.npair + pbd;
}

// GintS(i1, i2, i3, i4) (src/integrals.f90:371-493) with Kexn = 1: the
// two-electron sigma of the pair in both orders; 0 when a shell lies above
// LmaxS/NmaxS, the shell indices sum above Nsum or the l sum is odd.  The
// lookup order of the quadruple is GintS's (lowest shell first, its tie
// rules, and with Ksym = 0 nb <= nd); the multipoles k run over the given
// order's triangle rule up to Kmax; with Kdsig the value is extrapolated in
// E_k + Eps(a) + Eps(b) - E of the quadruple as given.
PCI_HD double gints(const Tables &T, int i1, int i2, int i3, int i4, double ek) {
#pragma clang fp contract(off)
    double e = 0.0;
    int is = 1, ia = i1, ib = i2, ic = i3, id = i4, t;
    for (int ii = 0; ii < 2; ii++) {
        const int ma = T.jz[ia], mb = T.jz[ib], mc = T.jz[ic], md = T.jz[id];
        int na = T.nh[ia], nb = T.nh[ib], nc = T.nh[ic], nd = T.nh[id];
        const int na0 = na, nb0 = nb;
        const int la = T.ll[na], lb = T.ll[nb], lc = T.ll[nc], ld = T.ll[nd];
        int ml = la > lb ? la : lb; ml = ml > lc ? ml : lc; ml = ml > ld ? ml : ld;
        int mn = na > nb ? na : nb; mn = mn > nc ? mn : nc; mn = mn > nd ? mn : nd;
        if (ml > T.lmaxs || mn > T.nmaxs || na + nb + nc + nd > T.nsum) return 0.0;
        if ((la + lb + lc + ld) & 1) return 0.0;
        const int ja = T.jj[na], jb = T.jj[nb], jc = T.jj[nc], jd = T.jj[nd];
        if (na > nc && nb >= nc && nd >= nc) { t = na; na = nc; nc = t; t = nb; nb = nd; nd = t; }
        else if (na > nb && nc > nb && nd >= nb) { t = na; na = nb; nb = t; t = nc; nc = nd; nd = t; }
        else if (na > nd && nb > nd && nc > nd) { t = na; na = nd; nd = t; t = nb; nb = nc; nc = t; }
        if (na == nb && nc > nd) { t = nc; nc = nd; nd = t; }
        if (na == nd && nb > nc) {
            if (na != nc) { t = nc; nc = nb; nb = t; }
            else { t = nd; nd = nb; nb = t; }
        }
        if (na == nc && nb > nd) { t = nb; nb = nd; nd = t; }
        if (T.ksym == 0) {
            if (nb > nd) { t = nb; nb = nd; nd = t; }
            if (na == nb && nc > nd) { t = nc; nc = nd; nd = t; }
        }
        const int iac0 = T.nx * (na - T.nso - 1) + (nc - T.nso);
        const int ibd0 = T.nx * (nb - T.nso - 1) + (nd - T.nso);
        int dac = ja - jc; dac = dac < 0 ? -dac : dac;
        int dbd = jb - jd; dbd = dbd < 0 ? -dbd : dbd;
        int kmn = dac / 2, kmx = (ja + jc) / 2;
        if (kmn < dbd / 2) kmn = dbd / 2;
        if (kmx > (jb + jd) / 2) kmx = (jb + jd) / 2;
        if (kmx > T.kmaxs) kmx = T.kmaxs;
        for (int k = kmn; k <= kmx; k++) {
            const int64_t key = ((int64_t)(T.nx * T.nx * k + iac0) << 32) | (uint32_t)ibd0;
            int64_t slot = s2_hash(key, T.s2mask);
            while (T.s2key[slot] != key && T.s2key[slot] != -1) slot = (slot + 1) & T.s2mask;
            if (T.s2key[slot] != key) {  // "GintS: missing integral"
                *T.err = 1;
                return __builtin_nan("");
            }
            const double *v = T.s2v + 3 * slot;
            double rabcd = v[0];
            if (rabcd != 0.0) {
                const double dabcd = v[1];
                if ((double)T.kdsig * dabcd != 0.0) {
                    const double de = ek + T.eps[na0] + T.eps[nb0] - v[2];
                    const double d1 = dabcd / rabcd;
                    double dd = d1 * de;
                    if (dd > 0.5) dd = 0.5;
                    if (d1 > 0.0) rabcd = rabcd / (1.0 - dd);   // normal variant
                    else rabcd = rabcd * (1.0 + dd);            // anomalous: two-side extrapolation
                }
                const double g1 = gaunt(T, k, ja, ma, jc, mc);
                const double g2 = gaunt(T, k, jd, md, jb, mb);
                e = e + (double)is * g1 * g2 * rabcd;
            }
        }
        t = ic; ic = id; id = t;
        is = -is;
    }
    return e;
}

// <i1 i2 || i3 i4>; ek = E_k, the Diag of the column determinant (GintS with
// Kdsig)
PCI_HD double gint(const Tables &T, int i1, int i2, int i3, int i4, double ek = 0.0) {
#pragma clang fp contract(off)
    const bool l_br = T.kbrt != 0;
    double e = 0.0;
    int is = 1, ia = i1, ib = i2, ic = i3, id = i4;
    for (int ii = 0; ii < 2; ii++) {
        const int ma = T.jz[ia], mb = T.jz[ib], mc = T.jz[ic], md = T.jz[id];
        int na = T.nh[ia], nb = T.nh[ib], nc = T.nh[ic], nd = T.nh[id];
        const int la = T.ll[na], lb = T.ll[nb], lc = T.ll[nc], ld = T.ll[nd];
        if (((la + lb + lc + ld) & 1) == 0) {
            const int ja = T.jj[na], jb = T.jj[nb], jc = T.jj[nc], jd = T.jj[nd];
            int ibr = 0, k;
            if (na > nc) { k = na; na = nc; nc = k; ibr++; }
            if (nb > nd) { k = nb; nb = nd; nd = k; ibr++; }
            if (na > nb) { k = na; na = nb; nb = k; k = nc; nc = nd; nd = k; }
            if (na == nb && nc > nd) { k = nc; nc = nd; nd = k; }
            ibr = (ibr == 1) ? -1 : 1;
            int dac = ja - jc; dac = dac < 0 ? -dac : dac;
            int dbd = jb - jd; dbd = dbd < 0 ? -dbd : dbd;
            int kmin = dac / 2, kmax = (ja + jc) / 2;
            if (kmin < dbd / 2) kmin = dbd / 2;
            if (kmax > (jb + jd) / 2) kmax = (jb + jd) / 2;
            int is_br = ((la + lc + kmin) & 1) ? ibr : 1;
            const int64_t q = quad_index(T, na, nc, nb, nd);
            const int64_t base = T.quad[q];
            for (k = kmin; k <= kmax; k++) {
                const bool l_pr = ((k + la + lc) & 1) == 0;
                if (l_pr || l_br) {
                    is_br = is_br * ibr;
                    double r = base >= 0 ? T.rk[base + k] : __builtin_nan("");
                    if (r != r) {
                        *T.err = 1;
                        return __builtin_nan("");
                    }
                    double rabcd = r;
                    if (l_br) rabcd = rabcd + is_br * T.rk2[base + k];
                    if (T.ksig >= 2) {
                        // screening when max(na, nb, nc, nd) > Nd or max(l) > Lmax:
                        // Fortran names ignore case, so that Nd is Gint's own
                        // local nd (the canonical d shell), and Lmax is a module
                        // variable pconf never sets (0 in a static build)
                        int mn = na > nb ? na : nb; mn = mn > nc ? mn : nc; mn = mn > nd ? mn : nd;
                        int ml = la > lb ? la : lb; ml = ml > lc ? ml : lc; ml = ml > ld ? ml : ld;
                        if ((mn > nd || ml > 0) && k < 10) rabcd = T.scr[k] * rabcd;
                    }
                    const double g1 = gaunt(T, k, ja, ma, jc, mc);
                    const double g2 = gaunt(T, k, jd, md, jb, mb);
                    e = e + (double)is * g1 * g2 * rabcd;
                }
            }
        }
        int t = ic; ic = id; id = t;
        is = -is;
    }
    if (T.ksig >= 2) e = e + gints(T, i1, i2, i3, i4, ek);
    return e;
}

PCI_HD double plj(const Tables &T, int ia, int ib) {
    double t = 0.0;
    const int na = T.nh[ia], nb = T.nh[ib];
    if (na == nb) {
        const int ma = T.jz[ia], mb = T.jz[ib];
        if (ma == mb + 2) {
            const int ja = T.jj[na];
            t = __builtin_sqrt((double)(ja * (ja + 2) - ma * mb));
        }
    }
    return t;
}

// Bit helpers over W words.
template <int W>
PCI_HD int rank_below(const uint64_t *d, int pos0) {   // set bits of d below 0-based bit pos0
    int w = pos0 >> 6, r = 0;
#pragma unroll
    for (int i = 0; i < W; i++) {
        if (i < w) r += popc64(d[i]);
        else if (i == w) r += popc64(d[i] & ((1ull << (pos0 & 63)) - 1ull));
    }
    return r;
}

template <int W>
PCI_HD void two_lowest(const uint64_t *m, int &lo, int &hi) {  // 0-based bit positions of the (up to) two set bits
    lo = -1; hi = -1;
#pragma unroll
    for (int i = 0; i < W; i++) {
        uint64_t x = m[i];
        while (x) {
            int b = ctz64(x) + 64 * i;
            x &= x - 1;
            if (lo < 0) lo = b; else if (hi < 0) hi = b;
        }
    }
}

// J^2 element F_J2 (src/formj2.f90:13-46) for nf = 0 (A == B) or nf = 2.
template <int W>
PCI_HD double fj2_diag(const Tables &T, const uint64_t *A) {
#pragma clang fp contract(off)
    double t = (double)(T.mj * T.mj);
#pragma unroll
    for (int i = 0; i < W; i++) {
        uint64_t x = A[i];
        while (x) {
            const int a = ctz64(x) + 64 * i + 1;
            x &= x - 1;
            const int na = T.nh[a], ja = T.jj[na], ma = T.jz[a];
            t = t + (double)(ja * (ja + 2)) - (double)(ma * ma);
            // pairs with later positions
            uint64_t y = x;
            for (int j = i; j < W; j++) {
                uint64_t z = (j == i) ? y : A[j];
                while (z) {
                    const int b = ctz64(z) + 64 * j + 1;
                    z &= z - 1;
                    const double p1 = plj(T, a, b), p2 = plj(T, b, a);
                    t = t - p1 * p1 - p2 * p2;
                }
            }
        }
    }
    return t;
}

PCI_HD double fj2_double(const Tables &T, int is, int ia, int ic, int ib, int id) {
#pragma clang fp contract(off)
    return is * (plj(T, ia, ic) * plj(T, id, ib) + plj(T, ic, ia) * plj(T, ib, id)
                 - plj(T, ia, id) * plj(T, ic, ib) - plj(T, id, ia) * plj(T, ib, ic));
}

// Matrix element <det1|H|det2> exactly as Hmltn(idet1, is, nf, ...) after
// Rspq_phase1/2 (src/pconf.F90:1244-1256).  nf: popcount(A^B)/2, must be <= 2.
template <int W>
PCI_HD double element(const Tables &T, const uint64_t *A, const uint64_t *B, int nf, double ek = 0.0) {
#pragma clang fp contract(off)
    double t = 0.0;
    if (nf == 0) {
#pragma unroll
        for (int i = 0; i < W; i++) {
            uint64_t x = A[i];
            while (x) {
                const int a = ctz64(x) + 64 * i + 1;
                x &= x - 1;
                uint64_t y = x;
                for (int j = i; j < W; j++) {
                    uint64_t z = (j == i) ? y : A[j];
                    while (z) {
                        const int b = ctz64(z) + 64 * j + 1;
                        z &= z - 1;
                        t = t + gint(T, a, b, a, b, ek) * 1;
                    }
                }
                t = t + hint(T, a, a, ek) * 1;
            }
        }
        if (T.gj != 0.0) t = t + T.gj * fj2_diag<W>(T, A);  // + 0*F_J2 leaves every nonzero t unchanged
        return t;
    }
    uint64_t hA[W], hB[W];
#pragma unroll
    for (int i = 0; i < W; i++) { hA[i] = A[i] & ~B[i]; hB[i] = B[i] & ~A[i]; }
    int alo, ahi, blo, bhi;
    two_lowest<W>(hA, alo, ahi);
    two_lowest<W>(hB, blo, bhi);
    if (nf == 1) {
        const int ra = rank_below<W>(A, alo) + 1, rb = rank_below<W>(B, blo) + 1;
        const int d = rb - ra;
        const int is = ((d < 0 ? -d : d) & 1) ? -1 : 1;
        const int a = alo + 1, b = blo + 1;
#pragma unroll
        for (int i = 0; i < W; i++) {
            uint64_t x = A[i];
            while (x) {
                const int c = ctz64(x) + 64 * i + 1;
                x &= x - 1;
                if (c != a) t = t + gint(T, b, c, a, c, ek) * is;
            }
        }
        t = t + hint(T, b, a, ek) * is;
        return t;
    }
    // nf == 2: pairs (a_hi, b_hi) and (a_lo, b_lo)
    const int ra2 = rank_below<W>(A, ahi) + 1, rb2 = rank_below<W>(B, bhi) + 1;
    const int ra3 = rank_below<W>(A, alo) + 1, rb3 = rank_below<W>(B, blo) + 1;
    int is = 1;
    int d = rb2 - ra2; if ((d < 0 ? -d : d) & 1) is = -is;
    d = rb3 - ra3;     if ((d < 0 ? -d : d) & 1) is = -is;
    const int i1 = alo + 1, j1 = ahi + 1, i2 = blo + 1, j2 = bhi + 1;
    // Hmltn(idet1, is, 2, i2=b_lo, i1=a_lo, j2=b_hi, j1=a_hi): Gint(i2, j2, i1, j1)
    t = t + gint(T, i2, j2, i1, j1, ek) * is;
    if (T.gj != 0.0) t = t + T.gj * fj2_double(T, is, i2, i1, j2, j1);
    return t;
}

// J^2 element of FormJ (src/formj2.f90:140-156): nf 0 or 2, else 0.
template <int W>
PCI_HD double j2_element(const Tables &T, const uint64_t *A, const uint64_t *B, int nf) {
    if (nf == 0) return fj2_diag<W>(T, A);
    if (nf != 2) return 0.0;
    uint64_t hA[W], hB[W];
#pragma unroll
    for (int i = 0; i < W; i++) { hA[i] = A[i] & ~B[i]; hB[i] = B[i] & ~A[i]; }
    int alo, ahi, blo, bhi;
    two_lowest<W>(hA, alo, ahi);
    two_lowest<W>(hB, blo, bhi);
    const int ra2 = rank_below<W>(A, ahi) + 1, rb2 = rank_below<W>(B, bhi) + 1;
    const int ra3 = rank_below<W>(A, alo) + 1, rb3 = rank_below<W>(B, blo) + 1;
    int is = 1;
    int d = rb2 - ra2; if ((d < 0 ? -d : d) & 1) is = -is;
    d = rb3 - ra3;     if ((d < 0 ? -d : d) & 1) is = -is;
    return fj2_double(T, is, blo + 1, alo + 1, bhi + 1, ahi + 1);
}

}  // namespace pci
