/*
 * CPU oracle for the pconf CI hot path.  TEST INFRASTRUCTURE ONLY: this file is
 * the checker that the MI355X kernels are compared against; it is never linked
 * into, called by, or used as a fallback for the product library.
 *
 * Each function restates one piece of the reference algorithm, in plain C,
 * following the cited lines of ud-pci/pCI.  Arithmetic order follows the
 * reference term by term and the file is compiled with -ffp-contract=off so
 * that matrix elements (and therefore pconf's exact-zero filtering) are
 * reproduced bit for bit.
 *
 *   Dinit                    src/determinants.f90:47-105
 *   Ndet (odometer)          src/determinants.f90:350-437
 *   Det_List + sort_det      src/determinants.f90:144-183, 811-836
 *   Rspq_phase1/2            src/determinants.f90:595-725
 *   CompD                    src/determinants.f90:748-779
 *   Gaunt (LUT + symmetry)   src/integrals.f90:509-561, 702-715
 *   Hint                     src/integrals.f90:122-151, 663-674
 *   Gint + FindGint          src/integrals.f90:153-293, 563-632
 *   Hmltn                    src/pconf.F90:1653-1692
 *   F_J2, Plj                src/formj2.f90:13-67
 *   FormH (npes = 1 loop)    src/pconf.F90:1226-1356, 1597-1606
 *   FormJ (npes = 1 loop)    src/formj2.f90:120-184
 *   J_av                     src/formj2.f90:355-387
 *   Jterm (J terms, Njt)     src/determinants.f90:185-348
 *   Mxmpy (one column)       src/davidson.f90:572-647
 *   nonequiv_conf            src/csf.f90:454-536
 *   build_j2                 src/csf.f90:409-452
 *   idif                     src/csf.f90:1189-1227
 *   formh_sym + compute_zzc_block + Hmatrix
 *                            src/csf.f90:538-1051
 *
 * Indices are 1-based like the reference: positions 1..Nst, shells 1..Ns,
 * determinants 1..Nd.  Iarr is stored determinant-major: iarr[(n-1)*Ne + i].
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    /* basis (1-based, index 0 unused) */
    int Ns, Nso, Nsu, Nx;
    int *Nn, *Ll, *Kk, *Jj, *Nf0;
    /* configuration list: entries 1..Nsp (Nip shell, Nq occupation) */
    int Nsp, Nc, Ne, Mj;
    int *Nip, *Nq, *Nq0, *Nvc, *Nc0;
    /* positions 1..Npos (all Ns shells), Nst = positions of shells 1..Nsu */
    int Npos, Nst;
    int *Jz, *Nh;
    /* determinants */
    int Nd;
    int *Ndc;   /* 1..Nc */
    int *iarr;  /* Nd*Ne */
    /* integrals */
    double *HintLUT; /* 1..Nx*Nx */
    int Kbrt;
    int64_t Ngint;
    double *Rint2;   /* Ngint*(1|2) */
    int *Iint2, *Iint3;
    int64_t cap;
    int *hIac, *hIbd;
    int64_t *hPos;
    /* gaunt */
    int Nlx;
    int gmax;
    double *GauntLUT; /* 0..gmax */
    double Gj;
    /* MBPT one-electron effective integrals (Ksig >= 1) */
    int Ksig, Kdsig, NmaxS, LmaxS;
    int *SgPos;                /* HintSPosLUT: 1..Nx*Nx -> 1..NhintS, 0 absent */
    double *Rsig, *Dsig, *Esig; /* 1..NhintS */
    double *Eps;               /* 1..Ns orbital energies (CONF.DAT) */
    double *DiagS;             /* 1..Nd: FormD, E_0 - sum of Eps (Kdsig != 0) */
    double E_k;                /* DiagS of the current column determinant */
    /* Ksig = 2: screening of Gint's R^k and the two-electron sigma (SCRC.CON) */
    int Ksym, KmaxS, Nsum;
    double Scr[10];
    int64_t NgintS, capS;
    double *Rint2S, *Dint2S, *Eint2S; /* 1..NgintS */
    int *sIac, *sIbd;
    int64_t *sPos;
} orc_ctx;

static void *xcalloc(size_t n, size_t s) { void *p = calloc(n ? n : 1, s); return p; }

orc_ctx *orc_create(void) { return (orc_ctx *)xcalloc(1, sizeof(orc_ctx)); }

void orc_free(orc_ctx *c) {
    if (!c) return;
    free(c->SgPos); free(c->Rsig); free(c->Dsig); free(c->Esig); free(c->Eps); free(c->DiagS);
    free(c->Rint2S); free(c->Dint2S); free(c->Eint2S); free(c->sIac); free(c->sIbd); free(c->sPos);
    free(c->Nn); free(c->Ll); free(c->Kk); free(c->Jj); free(c->Nf0);
    free(c->Nip); free(c->Nq); free(c->Nq0); free(c->Nvc); free(c->Nc0);
    free(c->Jz); free(c->Nh); free(c->Ndc); free(c->iarr);
    free(c->HintLUT); free(c->Rint2); free(c->Iint2); free(c->Iint3);
    free(c->hIac); free(c->hIbd); free(c->hPos); free(c->GauntLUT);
    free(c);
}

/* ---------------------------------------------------------------------------
 * Basis, configurations and Dinit (src/determinants.f90:47-105; src/pconf.F90:506-600)
 * shells given 0-based arrays of length Ns: n, l, kappa, 2j.
 * entries: Nsp configuration entries (shell index 1-based, occupation); the
 * first Nso entries are core shells.
 * ------------------------------------------------------------------------- */
int orc_set_space(orc_ctx *c, int Ns, int Nso, const int *nn, const int *ll, const int *kk,
                  const int *jj, int Nsp, const int *nip, const int *nq, int Ne, int Mj, double Gj) {
    c->Ns = Ns; c->Nso = Nso; c->Nsp = Nsp; c->Ne = Ne; c->Mj = Mj; c->Gj = Gj;
    c->Nn = xcalloc(Ns + 1, sizeof(int)); c->Ll = xcalloc(Ns + 1, sizeof(int));
    c->Kk = xcalloc(Ns + 1, sizeof(int)); c->Jj = xcalloc(Ns + 1, sizeof(int));
    c->Nf0 = xcalloc(Ns + 1, sizeof(int));
    for (int i = 0; i < Ns; i++) { c->Nn[i+1] = nn[i]; c->Ll[i+1] = ll[i]; c->Kk[i+1] = kk[i]; c->Jj[i+1] = jj[i]; }
    c->Nip = xcalloc(Nsp + 1, sizeof(int)); c->Nq = xcalloc(Nsp + 1, sizeof(int));
    c->Nq0 = xcalloc(Nsp + 1, sizeof(int));
    c->Nsu = 0;
    for (int i = 0; i < Nsp; i++) {
        c->Nip[i+1] = nip[i]; c->Nq[i+1] = nq[i];
        if (c->Nsu < nip[i]) c->Nsu = nip[i];
    }
    /* positions over all shells (Dinit) */
    int i0 = 0;
    for (int ni = 1; ni <= Ns; ni++) i0 += c->Jj[ni] + 1;
    c->Npos = i0;
    c->Jz = xcalloc(i0 + 1, sizeof(int)); c->Nh = xcalloc(i0 + 1, sizeof(int));
    i0 = 0;
    for (int ni = 1; ni <= Ns; ni++) {
        int nem = c->Jj[ni] + 1, imax = 2 * c->Jj[ni] + 1;
        c->Nf0[ni] = i0;
        for (int j = 1; j <= imax; j += 2) { i0++; c->Jz[i0] = j - nem; c->Nh[i0] = ni; }
    }
    c->Nst = 0;
    for (int ni = 1; ni <= c->Nsu; ni++) c->Nst += c->Jj[ni] + 1;
    /* configurations (src/pconf.F90:557-578, determinants.f90:84-103) */
    int maxc = Nsp + 1;
    c->Nvc = xcalloc(maxc + 1, sizeof(int)); c->Nc0 = xcalloc(maxc + 1, sizeof(int));
    int n = 0, ic = 0, i = 0; i0 = 0;
    for (int ni = Nso + 1; ni <= Nsp; ni++) {
        i++;
        c->Nq0[ni] = n;
        n += c->Nq[ni];
        if (n >= Ne) {
            if (n > Ne) return -1; /* wrong number of electrons */
            ic++;
            c->Nvc[ic] = i; c->Nc0[ic] = Nso + i0;
            i0 += i; n = 0; i = 0;
        }
    }
    c->Nc = ic;
    return ic;
}

/* Ndet: next determinant of configuration ic (src/determinants.f90:350-437).
 * idet is 1-based (idet[1..Ne]); returns M through *M. */
static void ndet(orc_ctx *c, int ic, int *fin, int *idet, int *M) {
    int n1 = c->Nc0[ic] + 1, n2 = c->Nc0[ic] + c->Nvc[ic];
    *M = 0;
    if (*fin) {
        for (int ni = n1; ni <= n2; ni++) {
            int nqi = c->Nq[ni];
            if (nqi) {
                int i = c->Nip[ni], n0 = c->Nq0[ni], if0 = c->Nf0[i] - n0;
                for (int iq = n0 + 1; iq <= n0 + nqi; iq++) idet[iq] = if0 + iq;
            }
        }
        *fin = 0;
    } else {
        int n3 = n1 + n2, found = 0;
        for (int is = n1; is <= n2 && !found; is++) {
            int ni = n3 - is, nqi = c->Nq[ni];
            if (!nqi) continue;
            int i = c->Nip[ni], n0 = c->Nq0[ni], i0 = n0 + 1, im = n0 + nqi;
            int n = c->Jj[i] + 1 + c->Nf0[i] - im, i1 = im + i0;
            for (int k = i0; k <= im; k++) {
                int iq = i1 - k;
                if (idet[iq] < n + iq) {
                    idet[iq]++;
                    int i2 = idet[iq] - iq;
                    for (int jq = iq + 1; jq <= im; jq++) idet[jq] = jq + i2;
                    for (int nj = ni + 1; nj <= n2; nj++) {
                        int nqj = c->Nq[nj];
                        if (nqj) {
                            int j = c->Nip[nj], m0 = c->Nq0[nj], jf0 = c->Nf0[j] - m0;
                            for (int jq = m0 + 1; jq <= m0 + nqj; jq++) idet[jq] = jq + jf0;
                        }
                    }
                    found = 1;
                    break;
                }
            }
        }
        if (!found) { *fin = 1; return; }
    }
    for (int iq = 1; iq <= c->Ne; iq++) *M += c->Jz[idet[iq]];
}


/* Jterm (src/determinants.f90:185-348): the doubled total-J values present
 * in the space (Jt, ascending) and the number of levels of each (Njt).  Per
 * configuration with determinants at M = Mj, the determinants with M >= Mj
 * are counted per M; walking M down from the top, the count left at each M
 * is the number of new J = M terms.  Returns Njd (or -1 if cap is too small). */
int orc_jterms(orc_ctx *c, int *jt, int64_t *njt, int cap) {
    int Ne = c->Ne, njd = 0, jmax = 0;
    int *idet = xcalloc(Ne + 1, sizeof(int));
    for (int ic = 1; ic <= c->Nc; ic++) {
        int fin = 1, M;
        ndet(c, ic, &fin, idet, &M);
        while (!fin) {
            int im = (M - c->Mj) / 2 + 1;
            if (im > jmax) jmax = im;
            ndet(c, ic, &fin, idet, &M);
        }
    }
    int64_t *nmj = xcalloc(jmax + 2, sizeof(int64_t));
    for (int ic = 1; ic <= c->Nc; ic++) {
        int fin = 1, M, ndi = 0, imax = 1;
        ndet(c, ic, &fin, idet, &M);
        while (!fin) {
            if (M >= c->Mj) {
                if (M == c->Mj) ndi++;
                int im = (M - c->Mj) / 2 + 1;
                if (im > imax) imax = im;
            }
            ndet(c, ic, &fin, idet, &M);
        }
        if (!ndi) continue;
        for (int im = 1; im <= imax; im++) nmj[im] = 0;
        fin = 1;
        ndet(c, ic, &fin, idet, &M);
        while (!fin) {
            if (M >= c->Mj) nmj[(M - c->Mj) / 2 + 1]++;
            ndet(c, ic, &fin, idet, &M);
        }
        for (int im = imax; im >= 1; im--) {
            int64_t n = nmj[im];
            if (!n) continue;
            int mt = (im - 1) * 2 + c->Mj;
            for (int j = 1; j <= im; j++) nmj[j] -= n;
            int j;
            for (j = 0; j < njd; j++)
                if (jt[j] == mt) { njt[j] += n; break; }
            if (j == njd) {
                if (njd >= cap) { free(idet); free(nmj); return -1; }
                jt[njd] = mt; njt[njd] = n; njd++;
            }
        }
    }
    /* ascending by J (bubble sort, as the reference) */
    for (int swapped = 1; swapped;) {
        swapped = 0;
        for (int j = 1; j < njd; j++)
            if (jt[j] < jt[j - 1]) {
                int t = jt[j]; jt[j] = jt[j - 1]; jt[j - 1] = t;
                int64_t u = njt[j]; njt[j] = njt[j - 1]; njt[j - 1] = u;
                swapped = 1;
            }
    }
    free(idet); free(nmj);
    return njd;
}

/* Det_Number + Det_List with sort_det: returns Nd, fills Ndc and Iarr. */
int64_t orc_enumerate(orc_ctx *c) {
    int Ne = c->Ne;
    int *idet = xcalloc(Ne + 1, sizeof(int)), *sdet = xcalloc(Ne + 1, sizeof(int));
    free(c->Ndc);
    c->Ndc = xcalloc(c->Nc + 1, sizeof(int));
    int64_t nd = 0;
    for (int pass = 0; pass < 2; pass++) {
        nd = 0;
        for (int ic = 1; ic <= c->Nc; ic++) {
            int fin = 1, M, ndi = 0;
            ndet(c, ic, &fin, idet, &M);
            while (!fin) {
                if (M == c->Mj) {
                    ndi++;
                    if (pass == 1) {
                        /* sort_det: selection sort ascending */
                        for (int q = 1; q <= Ne; q++) sdet[q] = idet[q];
                        for (int a = 1; a <= Ne - 1; a++) {
                            int k = a;
                            for (int b = a + 1; b <= Ne; b++) if (sdet[b] < sdet[k]) k = b;
                            if (k != a) { int t = sdet[k]; sdet[k] = sdet[a]; sdet[a] = t; }
                        }
                        int64_t d = nd + ndi - 1;
                        for (int q = 1; q <= Ne; q++) c->iarr[d * Ne + q - 1] = sdet[q];
                    }
                }
                ndet(c, ic, &fin, idet, &M);
            }
            c->Ndc[ic] = ndi;
            nd += ndi;
        }
        if (pass == 0) {
            free(c->iarr);
            c->iarr = xcalloc((size_t)nd * Ne, sizeof(int));
        }
    }
    c->Nd = (int)nd;
    free(idet); free(sdet);
    return nd;
}

int orc_nst(orc_ctx *c) { return c->Nst; }
int orc_npos(orc_ctx *c) { return c->Npos; }
int orc_nc(orc_ctx *c) { return c->Nc; }
int orc_nsu(orc_ctx *c) { return c->Nsu; }

void orc_get_space(orc_ctx *c, int *iarr, int *ndc, int *jz, int *nh) {
    if (iarr) memcpy(iarr, c->iarr, sizeof(int) * (size_t)c->Nd * c->Ne);
    if (ndc) memcpy(ndc, c->Ndc + 1, sizeof(int) * c->Nc);
    if (jz) memcpy(jz, c->Jz + 1, sizeof(int) * c->Npos);
    if (nh) memcpy(nh, c->Nh + 1, sizeof(int) * c->Npos);
}

/* Use a caller-supplied determinant list (e.g. one read from CONF.DET). */
int orc_set_dets(orc_ctx *c, int nd, const int *iarr, const int *ndc) {
    free(c->iarr); free(c->Ndc);
    c->Nd = nd;
    c->iarr = xcalloc((size_t)nd * c->Ne, sizeof(int));
    memcpy(c->iarr, iarr, sizeof(int) * (size_t)nd * c->Ne);
    c->Ndc = xcalloc(c->Nc + 1, sizeof(int));
    memcpy(c->Ndc + 1, ndc, sizeof(int) * c->Nc);
    return 0;
}

/* ---------------------------------------------------------------------------
 * Integral tables: BuildHintLUT, BuildGintHash, BuildGauntLUT
 * (src/integrals.f90:604-632, 663-674, 702-715)
 * ------------------------------------------------------------------------- */
int orc_set_integrals(orc_ctx *c, int nx, int nhint, const double *rint1, const int *iint1,
                      int64_t ngint, int kbrt, const double *rint2, const int *iint2,
                      const int *iint3) {
    c->Nx = nx; c->Kbrt = kbrt; c->Ngint = ngint;
    free(c->HintLUT);
    c->HintLUT = xcalloc((size_t)nx * nx + 1, sizeof(double));
    for (int i = 0; i < nhint; i++) {
        if (iint1[i] < 1 || iint1[i] > nx * nx) return -1;
        c->HintLUT[iint1[i]] = rint1[i];
    }
    int w = kbrt ? 2 : 1;
    free(c->Rint2); free(c->Iint2); free(c->Iint3);
    c->Rint2 = xcalloc((size_t)ngint * w, sizeof(double));
    memcpy(c->Rint2, rint2, sizeof(double) * (size_t)ngint * w);
    c->Iint2 = xcalloc(ngint, sizeof(int)); c->Iint3 = xcalloc(ngint, sizeof(int));
    memcpy(c->Iint2, iint2, sizeof(int) * ngint); memcpy(c->Iint3, iint3, sizeof(int) * ngint);
    int64_t cap = 4;
    while (cap < 2 * (ngint > 1 ? ngint : 1)) cap *= 2;
    c->cap = cap;
    free(c->hIac); free(c->hIbd); free(c->hPos);
    c->hIac = xcalloc(cap, sizeof(int)); c->hIbd = xcalloc(cap, sizeof(int));
    c->hPos = xcalloc(cap, sizeof(int64_t));
    for (int64_t i = 1; i <= ngint; i++) {
        int64_t h = (int64_t)iint2[i-1] * 104729 + (int64_t)iint3[i-1];
        int64_t slot = h % cap; if (slot < 0) slot += cap;
        while (c->hPos[slot] != 0) slot = (slot + 1) % cap;
        c->hIac[slot] = iint2[i-1]; c->hIbd[slot] = iint3[i-1]; c->hPos[slot] = i;
    }
    return 0;
}

int orc_set_gaunt(orc_ctx *c, int nlx, int ngaunt, const int *in, const double *gnt) {
    c->Nlx = nlx;
    int mx = 0;
    for (int i = 0; i < ngaunt; i++) if (in[i] > mx) mx = in[i];
    c->gmax = mx;
    free(c->GauntLUT);
    c->GauntLUT = xcalloc((size_t)mx + 1, sizeof(double));
    for (int i = 0; i < ngaunt; i++) c->GauntLUT[in[i]] = gnt[i];
    return 0;
}

static int64_t find_gint(const orc_ctx *c, int iac, int ibd) {
    int64_t h = (int64_t)iac * 104729 + (int64_t)ibd;
    int64_t slot = h % c->cap; if (slot < 0) slot += c->cap;
    while (c->hPos[slot] != 0) {
        if (c->hIac[slot] == iac && c->hIbd[slot] == ibd) return c->hPos[slot];
        slot = (slot + 1) % c->cap;
    }
    return 0;
}

/* Gaunt(k, j1, m1, j2, m2) with half-integer reals (src/integrals.f90:509-561) */
double orc_gaunt(const orc_ctx *c, int k, double j1, double m1, double j2, double m2) {
    int is = 1;
    double g = 0.0, x;
    int im = (int)fabs(m2 - m1);
    if (k >= im) {
        if (k == 0) {
            if (j1 == j2) return 1.0;
        } else {
            if (j2 < j1) {
                x = j2; j2 = j1; j1 = x;
                x = m2; m2 = m1; m1 = x;
                if (im % 2) is = -is;
            }
            if (m1 <= 0.0) {
                m1 = -m1; m2 = -m2;
                int ij = (int)(k + j1 - j2);
                if (ij % 2) is = -is;
            }
            int ib1 = 2 * c->Nlx + 1, ib2 = ib1 * ib1;
            int ind = (int)(ib2 * (ib2 * k + 2 * (ib1 * j1 + j2)) + ib1 * (j1 + m1) + (j2 + m2));
            if (ind >= 0 && ind <= c->gmax) {
                g = c->GauntLUT[ind];
                return g * is;
            }
        }
    }
    return g * is;
}

/* HintS(na, nb, n0) (src/integrals.f90:321-369) with Kexn = 1 (the value
 * RintS sets, src/pconf.F90:700): the sigma correction of the pair na <= nb;
 * with Kdsig != 0 and D != 0 it is extrapolated in the energy
 * E_k + Eps(n0) - Esig.  Sets g_missing_sig when the pair is absent. */
static int g_missing_sig = 0;
static double orc_hints(const orc_ctx *c, int na, int nb, int n0) {
    double e = 0.0;
    if (c->Ll[na] > c->LmaxS || nb > c->NmaxS) return e;
    int ind = c->Nx * (na - c->Nso - 1) + (nb - c->Nso);
    int i = c->SgPos[ind];
    if (i == 0) { g_missing_sig = 1; return NAN; }
    e = c->Rsig[i];
    double d = c->Dsig[i];
    if (c->Kdsig * d != 0.0) {
        double de = (c->E_k + c->Eps[n0] - c->Esig[i]);
        double d1 = d / e;
        double dd = d1 * de;
        if (dd > 0.5) dd = 0.5;
        if (d1 > 0.0) e = e / (1.0 - dd);   /* normal variant */
        else e = e * (1.0 + dd);            /* anomalous, two-side extrapolation */
    }
    return e;
}

/* Hint(ia, ib) (src/integrals.f90:122-151), C_is = 0; + HintS with Ksig */
double orc_hint(const orc_ctx *c, int ia, int ib) {
    double e = 0.0;
    if (c->Jz[ia] == c->Jz[ib]) {
        int na = c->Nh[ia], nb = c->Nh[ib];
        int n0 = na;
        if (c->Kk[na] == c->Kk[nb]) {
            if (na > nb) { int t = na; na = nb; nb = t; }
            int ind = c->Nx * (na - c->Nso - 1) + (nb - c->Nso);
            e = c->HintLUT[ind];
            if (c->Ksig != 0) e = e + orc_hints(c, na, nb, n0);
        }
    }
    return e;
}

/* RintS for SGC.CON (src/pconf.F90:715-762) and BuildHintSLUT
 * (src/integrals.f90:676-687): entries (iint1s, rsig, dsig, esig) as read,
 * eps(1..Ns) from CONF.DAT; FormD (src/pconf.F90:852-880) when
 * Ksig * Kdsig != 0, which needs the determinants (call after enumerate). */
int orc_set_sigma(orc_ctx *c, int ksig, int kdsig, int nmaxs, int lmaxs, int nhints, const int *iint1s,
                  const double *rsig, const double *dsig, const double *esig, const double *eps, double e0) {
    c->Ksig = ksig; c->Kdsig = kdsig; c->NmaxS = nmaxs; c->LmaxS = lmaxs;
    free(c->SgPos); free(c->Rsig); free(c->Dsig); free(c->Esig); free(c->Eps); free(c->DiagS);
    c->SgPos = xcalloc((size_t)c->Nx * c->Nx + 1, sizeof(int));
    c->Rsig = xcalloc(nhints + 1, sizeof(double));
    c->Dsig = xcalloc(nhints + 1, sizeof(double));
    c->Esig = xcalloc(nhints + 1, sizeof(double));
    c->Eps = xcalloc(c->Ns + 1, sizeof(double));
    for (int i = 1; i <= nhints; i++) {
        int ind = iint1s[i - 1];
        if (ind < 1 || ind > c->Nx * c->Nx) return -1;
        c->SgPos[ind] = i;
        c->Rsig[i] = rsig[i - 1]; c->Dsig[i] = dsig[i - 1]; c->Esig[i] = esig[i - 1];
    }
    for (int ni = 1; ni <= c->Ns; ni++) c->Eps[ni] = eps[ni - 1];
    c->DiagS = xcalloc((size_t)c->Nd + 1, sizeof(double));
    if (ksig * kdsig != 0) {
        for (int n = 1; n <= c->Nd; n++) {
            double x = 0.0;
            for (int i = 0; i < c->Ne; i++) x = x + c->Eps[c->Nh[c->iarr[(size_t)(n - 1) * c->Ne + i]]];
            c->DiagS[n] = e0 - x;
        }
    }
    c->E_k = 0.0;
    return 0;
}
int orc_missing_sigma(void) { int m = g_missing_sig; g_missing_sig = 0; return m; }

/* RintS for Ksig = 2 (src/pconf.F90:764-840): the ten screening coefficients
 * of SGC.CON and SCRC.CON's Ksym, Kmax, Nsum (4 Nav or 4 NmaxS) and the
 * two-electron sigma integrals (R, D, E with their Iint2S/Iint3S keys), looked
 * up by an open-addressing table of the keys (any exact lookup gives the
 * reference's values; BuildGintSHash, src/integrals.f90:634-661, is one). */
int orc_set_sigma2(orc_ctx *c, int ksym, int kmax, int nsum, const double *scr, int64_t ngints, const int *iint2s,
                   const int *iint3s, const double *r, const double *d, const double *e) {
    c->Ksig = 2; c->Ksym = ksym; c->KmaxS = kmax; c->Nsum = nsum;
    for (int k = 0; k < 10; k++) c->Scr[k] = scr[k];
    free(c->Rint2S); free(c->Dint2S); free(c->Eint2S); free(c->sIac); free(c->sIbd); free(c->sPos);
    c->NgintS = ngints;
    c->Rint2S = xcalloc(ngints + 1, sizeof(double));
    c->Dint2S = xcalloc(ngints + 1, sizeof(double));
    c->Eint2S = xcalloc(ngints + 1, sizeof(double));
    int64_t cap = 4;
    while (cap < 2 * (ngints > 1 ? ngints : 1)) cap *= 2;
    c->capS = cap;
    c->sIac = xcalloc(cap, sizeof(int));
    c->sIbd = xcalloc(cap, sizeof(int));
    c->sPos = xcalloc(cap, sizeof(int64_t));
    for (int64_t i = 1; i <= ngints; i++) {
        c->Rint2S[i] = r[i - 1]; c->Dint2S[i] = d[i - 1]; c->Eint2S[i] = e[i - 1];
        int64_t h = ((int64_t)iint2s[i - 1] * 1000003 + iint3s[i - 1]) % cap;
        if (h < 0) h += cap;
        while (c->sPos[h] != 0) h = (h + 1) % cap;
        c->sIac[h] = iint2s[i - 1]; c->sIbd[h] = iint3s[i - 1]; c->sPos[h] = i;
    }
    return 0;
}

static int64_t find_gints(const orc_ctx *c, int iac, int ibd) {
    int64_t h = ((int64_t)iac * 1000003 + ibd) % c->capS;
    if (h < 0) h += c->capS;
    while (c->sPos[h] != 0) {
        if (c->sIac[h] == iac && c->sIbd[h] == ibd) return c->sPos[h];
        h = (h + 1) % c->capS;
    }
    return 0;
}
double orc_diags(const orc_ctx *c, int n) { return c->DiagS ? c->DiagS[n] : 0.0; }

static int g_missing = 0;
/* Lmax of Gint's screening test: a module variable (src/conf_variables.F90:12)
 * that pconf never assigns, so zero in a static build */
#define REF_LMAX 0

/* GintS(i1,i2,i3,i4) (src/integrals.f90:371-493), Kexn = 1: the two-electron
 * sigma of the pair, both orders (direct, then c <-> d with the opposite
 * sign); 0 when a shell lies above LmaxS/NmaxS, the shell indices sum above
 * Nsum or the l sum is odd.  Keys from the quadruple in GintS's canonical
 * order (lowest shell first; ties and Ksym = 0 as the reference breaks them),
 * multipoles k from the given order's triangle rule up to Kmax.  With Kdsig
 * the value is extrapolated in E_k + Eps(a) + Eps(b) - E of the quadruple as
 * given. */
static double orc_gints(const orc_ctx *c, int i1, int i2, int i3, int i4) {
    double e = 0.0;
    int is = 1, ia = i1, ib = i2, ic = i3, id = i4, t;
    for (int ii = 1; ii <= 2; ii++) {
        int ma = c->Jz[ia], mb = c->Jz[ib], mc = c->Jz[ic], md = c->Jz[id];
        int na = c->Nh[ia], nb = c->Nh[ib], nc = c->Nh[ic], nd = c->Nh[id];
        int na0 = na, nb0 = nb;
        int la = c->Ll[na], lb = c->Ll[nb], lc = c->Ll[nc], ld = c->Ll[nd];
        int ml = la > lb ? la : lb; ml = ml > lc ? ml : lc; ml = ml > ld ? ml : ld;
        int mn = na > nb ? na : nb; mn = mn > nc ? mn : nc; mn = mn > nd ? mn : nd;
        if (ml > c->LmaxS || mn > c->NmaxS || na + nb + nc + nd > c->Nsum) return 0.0;
        if ((la + lb + lc + ld) % 2) return 0.0;
        int ja = c->Jj[na], jb = c->Jj[nb], jc = c->Jj[nc], jd = c->Jj[nd];
        if (na > nc && nb >= nc && nd >= nc) { t = na; na = nc; nc = t; t = nb; nb = nd; nd = t; }
        else if (na > nb && nc > nb && nd >= nb) { t = na; na = nb; nb = t; t = nc; nc = nd; nd = t; }
        else if (na > nd && nb > nd && nc > nd) { t = na; na = nd; nd = t; t = nb; nb = nc; nc = t; }
        if (na == nb && nc > nd) { t = nc; nc = nd; nd = t; }
        if (na == nd && nb > nc) {
            if (na != nc) { t = nc; nc = nb; nb = t; }
            else { t = nd; nd = nb; nb = t; }
        }
        if (na == nc && nb > nd) { t = nb; nb = nd; nd = t; }
        if (c->Ksym == 0) {
            if (nb > nd) { t = nb; nb = nd; nd = t; }
            if (na == nb && nc > nd) { t = nc; nc = nd; nd = t; }
        }
        int iac0 = c->Nx * (na - c->Nso - 1) + (nc - c->Nso);
        int ibd0 = c->Nx * (nb - c->Nso - 1) + (nd - c->Nso);
        int kmn = abs(ja - jc) / 2 + 1, k2 = abs(jb - jd) / 2 + 1;
        if (kmn < k2) kmn = k2;
        int kmx = (ja + jc) / 2 + 1; k2 = (jb + jd) / 2 + 1;
        if (kmx > k2) kmx = k2;
        if (kmx > c->KmaxS + 1) kmx = c->KmaxS + 1;
        for (int k1 = kmn; k1 <= kmx; k1++) {
            int k = k1 - 1;
            int64_t i = find_gints(c, c->Nx * c->Nx * k + iac0, ibd0);
            if (i == 0) { g_missing = 3; return NAN; }   /* "GintS: missing integral" */
            double rabcd = c->Rint2S[i];
            if (rabcd != 0.0) {
                double dabcd = c->Dint2S[i];
                if (c->Kdsig * dabcd != 0.0) {
                    double eabcd = c->Eint2S[i];
                    double de = c->E_k + c->Eps[na0] + c->Eps[nb0] - eabcd;
                    double d1 = dabcd / rabcd;
                    double dd = d1 * de;
                    if (dd > 0.5) dd = 0.5;
                    if (d1 > 0.0) rabcd = rabcd / (1.0 - dd);   /* normal variant */
                    else rabcd = rabcd * (1.0 + dd);            /* anomalous, two-side */
                }
                e = e + (double)is
                      * orc_gaunt(c, k, ja * 0.5, ma * 0.5, jc * 0.5, mc * 0.5)
                      * orc_gaunt(c, k, jd * 0.5, md * 0.5, jb * 0.5, mb * 0.5)
                      * rabcd;
            }
        }
        t = ic; ic = id; id = t;
        is = -is;
    }
    return e;
}

/* Gint(i1,i2,i3,i4) = <i1 i2 || i3 i4> (src/integrals.f90:153-293), K_is = 0.
 * Returns NAN (and sets *missing) when an integral is absent from CONF.INT. */
double orc_gint(const orc_ctx *c, int i1, int i2, int i3, int i4) {
    int l_br = c->Kbrt != 0;
    double e = 0.0;
    int is = 1, ia = i1, ib = i2, ic = i3, id = i4;
    for (int ii = 1; ii <= 2; ii++) {
        int ma = c->Jz[ia], mb = c->Jz[ib], mc = c->Jz[ic], md = c->Jz[id];
        if (ma + mb != mc + md) { g_missing = 2; return NAN; }
        int na = c->Nh[ia], nb = c->Nh[ib], nc = c->Nh[ic], nd = c->Nh[id];
        int la = c->Ll[na], lb = c->Ll[nb], lc = c->Ll[nc], ld = c->Ll[nd];
        int i = la + lb + lc + ld, ibr = 0, k;
        if (i % 2 == 0) {
            int ja = c->Jj[na], jb = c->Jj[nb], jc = c->Jj[nc], jd = c->Jj[nd];
            if (na > nc) { k = na; na = nc; nc = k; ibr++; }
            if (nb > nd) { k = nb; nb = nd; nd = k; ibr++; }
            if (na > nb) { k = na; na = nb; nb = k; k = nc; nc = nd; nd = k; }
            if (na == nb && nc > nd) { k = nc; nc = nd; nd = k; }
            ibr = (ibr == 1) ? -1 : 1;
            int iac0 = c->Nx * (na - c->Nso - 1) + (nc - c->Nso);
            int ibd0 = c->Nx * (nb - c->Nso - 1) + (nd - c->Nso);
            int kmin = abs(ja - jc) / 2 + 1; k = abs(jb - jd) / 2 + 1; if (kmin < k) kmin = k;
            int kmax = (ja + jc) / 2 + 1; k = (jb + jd) / 2 + 1; if (kmax > k) kmax = k;
            int i_br = la + lc + kmin - 1;
            int is_br = (i_br % 2) ? ibr : 1;
            for (int k1 = kmin; k1 <= kmax; k1++) {
                k = k1 - 1;
                int ipr = k + la + lc;
                int l_pr = (ipr % 2) == 0;
                if (l_pr || l_br) {
                    is_br = is_br * ibr;
                    int iac = c->Nx * c->Nx * k + iac0, ibd = ibd0;
                    int64_t mi8 = find_gint(c, iac, ibd);
                    if (mi8 == 0) { g_missing = 1; return NAN; }
                    double rabcd = c->Rint2[(l_br ? 2 : 1) * (mi8 - 1)];
                    if (l_br) rabcd = rabcd + is_br * c->Rint2[2 * (mi8 - 1) + 1];
                    if (c->Ksig >= 2) {
                        /* screening of the Coulomb (and Breit) integral when
                         * max(na, nb, nc, nd) > Nd or max(l) > Lmax: Fortran
                         * names ignore case, so that Nd is Gint's own local
                         * nd, the canonical d shell, and Lmax a module
                         * variable pconf never sets (0 in a static build):
                         * REF_LMAX */
                        int mn = na > nb ? na : nb; mn = mn > nc ? mn : nc; mn = mn > nd ? mn : nd;
                        int ml = la > lb ? la : lb; ml = ml > lc ? ml : lc; ml = ml > ld ? ml : ld;
                        if (mn > nd || ml > REF_LMAX) {
                            if (k < 10) rabcd = c->Scr[k] * rabcd;
                        }
                    }
                    e = e + (double)is
                          * orc_gaunt(c, k, ja * 0.5, ma * 0.5, jc * 0.5, mc * 0.5)
                          * orc_gaunt(c, k, jd * 0.5, md * 0.5, jb * 0.5, mb * 0.5)
                          * rabcd;
                }
            }
        }
        k = ic; ic = id; id = k;
        is = -is;
    }
    if (c->Ksig >= 2) e = e + orc_gints(c, i1, i2, i3, i4);
    return e;
}

int orc_missing(void) { int m = g_missing; g_missing = 0; return m; }

/* Plj, F_J2 (src/formj2.f90:13-67) */
static double plj(const orc_ctx *c, int ia, int ib) {
    double t = 0.0;
    int na = c->Nh[ia], nb = c->Nh[ib];
    if (na == nb) {
        int ma = c->Jz[ia], mb = c->Jz[ib];
        if (ma == mb + 2) { int ja = c->Jj[na]; t = sqrt((double)(ja * (ja + 2) - ma * mb)); }
    }
    return t;
}

double orc_fj2(const orc_ctx *c, const int *idet, int is, int nf, int ia, int ic, int ib, int id) {
    double t = 0.0;
    if (nf == 0) {
        t = (double)(c->Mj * c->Mj);
        for (int iq = 0; iq < c->Ne; iq++) {
            int a = idet[iq], na = c->Nh[a], ja = c->Jj[na], ma = c->Jz[a];
            t = t + (double)(ja * (ja + 2)) - (double)(ma * ma);
            for (int jq = iq + 1; jq < c->Ne; jq++) {
                int b = idet[jq];
                double p1 = plj(c, a, b), p2 = plj(c, b, a);
                t = t - p1 * p1 - p2 * p2;
            }
        }
    } else if (nf == 2) {
        t = is * (plj(c, ia, ic) * plj(c, id, ib) + plj(c, ic, ia) * plj(c, ib, id)
                  - plj(c, ia, id) * plj(c, ic, ib) - plj(c, id, ia) * plj(c, ib, ic));
    }
    return t;
}

/* Rspq_phase1 + Rspq_phase2 (src/determinants.f90:595-725).
 * id1, id2: 0-based C arrays of Ne sorted positions.
 * out: nf, is, i2, i3, j2, j3 where i2/i3 (j2/j3) are the differing positions
 * of det1 (det2) after phase 2, i2 the later one. */
void orc_rspq(int Ne, const int *id1, const int *id2, int *out) {
    int is = 1, nf = 3;
    int i[3] = {1, 0, 0}, j[3] = {1, 0, 0};
    int ni = 0, nj = 0;
    while (i[0] <= Ne && j[0] <= Ne) {
        int l1 = id1[i[0] - 1], l2 = id2[j[0] - 1];
        if (l1 == l2) { i[0]++; j[0]++; }
        else if (l1 > l2) {
            nj++;
            if (nj > 2) goto done;
            j[2] = j[1]; j[1] = j[0]; j[0]++;
        } else {
            ni++;
            if (ni > 2) goto done;
            i[2] = i[1]; i[1] = i[0]; i[0]++;
        }
    }
    if (i[0] > j[0]) {
        nf = ni;
        switch (Ne - j[0]) {
            case 0: j[2] = j[1]; j[1] = Ne; break;
            case 1: j[2] = Ne - 1; j[1] = Ne; break;
        }
    } else {
        nf = nj;
        switch (Ne - i[0]) {
            case 0: i[2] = i[1]; i[1] = Ne; break;
            case 1: i[2] = Ne - 1; i[1] = Ne; break;
        }
    }
    if (nf == 1) {
        int k = abs(j[1] - i[1]);
        if (k % 2 == 1) is = -is;
        i[1] = id1[i[1] - 1]; j[1] = id2[j[1] - 1];
    } else if (nf == 2) {
        int k = abs(j[1] - i[1]);
        if (k % 2 == 1) is = -is;
        k = abs(j[2] - i[2]);
        if (k % 2 == 1) is = -is;
        i[1] = id1[i[1] - 1]; j[1] = id2[j[1] - 1];
        i[2] = id1[i[2] - 1]; j[2] = id2[j[2] - 1];
    }
done:
    out[0] = nf; out[1] = is; out[2] = i[1]; out[3] = i[2]; out[4] = j[1]; out[5] = j[2];
}

/* CompD on shell lists (src/determinants.f90:748-779) */
static int compd(int Ne, const int *a, const int *b) {
    int ni = 0, nj = 0, i = 0, j = 0;
    while (i < Ne && j < Ne) {
        if (a[i] == b[j]) { i++; j++; }
        else if (a[i] > b[j]) { nj++; if (nj > 2) return 3; j++; }
        else { ni++; if (ni > 2) return 3; i++; }
    }
    return ni > nj ? ni : nj;
}

/* Hmltn(idet, is, nf, i2, i1, j2, j1) (src/pconf.F90:1653-1692); with_gj = 0
 * drops the Gj*F_J2 terms, which is csf.f90's Hmatrix (src/csf.f90:1008-1051,
 * kCSF /= 0). */
static double hmltn_core(const orc_ctx *c, const int *idet, int is, int nf, int i2, int i1, int j2, int j1,
                         int with_gj) {
    double t = 0.0;
    int Ne = c->Ne;
    if (nf == 2) {
        t = t + orc_gint(c, i2, j2, i1, j1) * is;
        if (with_gj) t = t + c->Gj * orc_fj2(c, idet, is, nf, i2, i1, j2, j1);
    } else if (nf == 1) {
        for (int iq = 0; iq < Ne; iq++) {
            int a = idet[iq];
            if (a != j1) t = t + orc_gint(c, j2, a, j1, a) * is;
        }
        t = t + orc_hint(c, j2, j1) * is;
    } else if (nf == 0) {
        for (int iq = 0; iq < Ne; iq++) {
            int a = idet[iq];
            for (int jq = iq + 1; jq < Ne; jq++) {
                int b = idet[jq];
                t = t + orc_gint(c, a, b, a, b) * is;
            }
            t = t + orc_hint(c, a, a) * is;
        }
        if (with_gj) t = t + c->Gj * orc_fj2(c, idet, is, nf, i2, i1, j2, j1);
    }
    return t;
}

double orc_hmltn(const orc_ctx *c, const int *idet, int is, int nf, int i2, int i1, int j2, int j1) {
    return hmltn_core(c, idet, is, nf, i2, i1, j2, j1, 1);
}

/* Element between determinants n and k (1-based); returns nf through *nfo. */
double orc_element(const orc_ctx *c, int n, int k, int *nfo) {
    int out[6];
    const int *d1 = c->iarr + (size_t)(n - 1) * c->Ne, *d2 = c->iarr + (size_t)(k - 1) * c->Ne;
    orc_rspq(c->Ne, d1, d2, out);
    if (nfo) *nfo = out[0];
    if (out[0] > 2) return 0.0;
    /* E_k = Diag(kk), the column determinant (src/pconf.F90:1275) */
    if (c->Kdsig != 0) ((orc_ctx *)c)->E_k = c->DiagS[k];
    /* Hmltn(idet1, iSign, diff, jIdx(3), iIdx(3), jIdx(2), iIdx(2)) */
    return orc_hmltn(c, d1, out[1], out[0], out[5], out[3], out[4], out[2]);
}

/* ---------------------------------------------------------------------------
 * FormH, npes = 1 (src/pconf.F90:1226-1356): rows r0..r1 (1-based inclusive).
 * Pass col == NULL to count only.  counts[n-r0] receives the number of stored
 * elements of row n.  Returns the number of stored elements; *numzero gets
 * the number of pairs with diff <= 2 whose element is exactly 0.0;
 * *ncand the number of pairs that reached Rspq.
 *
 * formh_rows is the row loop: colp/valp NULL counts only; cap NULL means the
 * caller's arrays hold every element; otherwise the arrays were malloc'd
 * with *cap entries and are doubled when full, as the reference's
 * vaccumulator grows its COO arrays (src/vaccumulator.f90:14-180).
 * ------------------------------------------------------------------------- */
static int64_t formh_rows(const orc_ctx *c, int r0, int r1, int64_t *counts, int **colp, double **valp,
                          int64_t *cap, int64_t *numzero, int64_t *ncand) {
    int Ne = c->Ne;
    int *cf1 = xcalloc(Ne, sizeof(int)), *cf2 = xcalloc(Ne, sizeof(int));
    int64_t pos = 0, nz = 0, nc = 0;
    for (int n = r0; n <= r1; n++) {
        const int *d1 = c->iarr + (size_t)(n - 1) * Ne;
        for (int q = 0; q < Ne; q++) cf1[q] = c->Nh[d1[q]];
        int64_t rowcnt = 0;
        int k = 0;
        for (int ic = 1; ic <= c->Nc; ic++) {
            int kx = c->Ndc[ic];
            if (k + kx > n) kx = n - k;
            if (kx == 0) continue;
            const int *d2 = c->iarr + (size_t)k * Ne;
            for (int q = 0; q < Ne; q++) cf2[q] = c->Nh[d2[q]];
            if (compd(Ne, cf1, cf2) > 2) { k += kx; continue; }
            for (int k1 = 1; k1 <= kx; k1++) {
                k++;
                int out[6];
                nc++;
                orc_rspq(Ne, d1, c->iarr + (size_t)(k - 1) * Ne, out);
                if (out[0] <= 2) {
                    if (c->Kdsig != 0) ((orc_ctx *)c)->E_k = c->DiagS[k];  /* src/pconf.F90:1275 */
                    double tt = orc_hmltn(c, d1, out[1], out[0], out[5], out[3], out[4], out[2]);
                    if (tt != 0.0) {
                        if (colp) {
                            if (cap && pos == *cap) {
                                *cap *= 2;
                                *colp = realloc(*colp, (size_t)*cap * sizeof(int));
                                *valp = realloc(*valp, (size_t)*cap * sizeof(double));
                            }
                            (*colp)[pos] = k;
                            (*valp)[pos] = tt;
                        }
                        pos++; rowcnt++;
                    } else {
                        nz++;
                    }
                }
            }
        }
        if (counts) counts[n - r0] = rowcnt;
    }
    free(cf1); free(cf2);
    if (numzero) *numzero = nz;
    if (ncand) *ncand = nc;
    return pos;
}

int64_t orc_formh(const orc_ctx *c, int r0, int r1, int64_t *counts, int *col, double *val,
                  int64_t *numzero, int64_t *ncand) {
    return formh_rows(c, r0, r1, counts, col ? &col : NULL, val ? &val : NULL, NULL, numzero, ncand);
}

/* FormH of rows r0..r1 in one pass (the reference evaluates each element
 * once and appends it): *colp and *valp are allocated here, grown as needed,
 * and released with orc_free_buf. */
int64_t orc_formh_grow(const orc_ctx *c, int r0, int r1, int64_t *counts, int **colp, double **valp,
                       int64_t *numzero, int64_t *ncand) {
    int64_t cap = 1 << 16;
    *colp = malloc((size_t)cap * sizeof(int));
    *valp = malloc((size_t)cap * sizeof(double));
    return formh_rows(c, r0, r1, counts, colp, valp, &cap, numzero, ncand);
}

void orc_free_buf(void *p) { free(p); }

/* FormJ, npes = 1 (src/formj2.f90:120-184): J^2 within configuration blocks. */
int64_t orc_formj(const orc_ctx *c, int64_t *counts, int *col, double *val) {
    int Ne = c->Ne;
    int64_t pos = 0;
    int n = 0;
    for (int ic1 = 1; ic1 <= c->Nc; ic1++) {
        int ndn = c->Ndc[ic1];
        for (int n1 = 1; n1 <= ndn; n1++) {
            n++;
            const int *d1 = c->iarr + (size_t)(n - 1) * Ne;
            int k = n - n1;
            int64_t rowcnt = 0;
            for (int k1 = 1; k1 <= n1; k1++) {
                k++;
                int out[6];
                orc_rspq(Ne, d1, c->iarr + (size_t)(k - 1) * Ne, out);
                if (out[0] == 0 || out[0] == 2) {
                    double tt = orc_fj2(c, d1, out[1], out[0], out[5], out[3], out[4], out[2]);
                    if (tt != 0.0) {
                        if (col) { col[pos] = k; val[pos] = tt; }
                        pos++; rowcnt++;
                    }
                }
            }
            if (counts) counts[n - 1] = rowcnt;
        }
    }
    return pos;
}

/* J_av (src/formj2.f90:355-387) over rows/columns <= nx; returns J. */
double orc_jav(int nd, const int64_t *rowptr, const int *col, const double *val, const double *x, int nx) {
    double xj = 0.0;
    for (int n = 1; n <= nd; n++) {
        if (n > nx) continue;
        for (int64_t l = rowptr[n - 1]; l < rowptr[n]; l++) {
            int k = col[l];
            double t = val[l];
            if (k <= nx) {
                double r = t * x[k - 1] * x[n - 1];
                if (n != k) r = r + r;
                xj = xj + r;
            }
        }
    }
    return 0.5 * (sqrt(1.0 + xj) - 1.0);
}

/* Mxmpy for one column (src/davidson.f90:600-637): y = (H - hmin I) x by
 * symmetric scatter over the stored lower triangle.  diag (optional) receives
 * the shifted diagonal. */
void orc_mxmpy(int nd, const int64_t *rowptr, const int *col, const double *val, double hmin,
               const double *x, double *y, double *diag) {
    for (int n = 0; n < nd; n++) y[n] = 0.0;
    for (int n = 1; n <= nd; n++) {
        for (int64_t l = rowptr[n - 1]; l < rowptr[n]; l++) {
            int k = col[l];
            double t = val[l];
            if (n == k) {
                t = t - hmin;
                if (diag) diag[n - 1] = t;
            }
            y[n - 1] = y[n - 1] + t * x[k - 1];
            if (n != k) y[k - 1] = y[k - 1] + t * x[n - 1];
        }
    }
}

/* Mxmpy over the rows one MPI rank owns (src/davidson.f90:617-634): rows
 * r0 .. r0+nrows-1 (1-based) of the lower-triangle CSR (rowptr[nrows+1]
 * from 0, 1-based columns), scattered symmetrically into the full-length y,
 * which is cleared first (before the Allreduce that sums the ranks' y). */
void orc_mxmpy_rows(int nd, int r0, int nrows, const int64_t *rowptr, const int *col, const double *val,
                    double hmin, const double *x, double *y) {
    for (int n = 0; n < nd; n++) y[n] = 0.0;
    for (int i = 0; i < nrows; i++) {
        const int n = r0 + i;
        for (int64_t l = rowptr[i]; l < rowptr[i + 1]; l++) {
            int k = col[l];
            double t = val[l];
            if (n == k) t = t - hmin;
            y[n - 1] = y[n - 1] + t * x[k - 1];
            if (n != k) y[k - 1] = y[k - 1] + t * x[n - 1];
        }
    }
}

/* Mxmpy over an arbitrary set of rows one rank holds: local row i is global
 * row rows[i] (1-based); otherwise as orc_mxmpy_rows. */
void orc_mxmpy_list(int nd, int nrows, const int *rows, const int64_t *rowptr, const int *col,
                    const double *val, double hmin, const double *x, double *y) {
    for (int n = 0; n < nd; n++) y[n] = 0.0;
    for (int i = 0; i < nrows; i++) {
        const int n = rows[i];
        for (int64_t l = rowptr[i]; l < rowptr[i + 1]; l++) {
            int k = col[l];
            double t = val[l];
            if (n == k) t = t - hmin;
            y[n - 1] = y[n - 1] + t * x[k - 1];
            if (n != k) y[k - 1] = y[k - 1] + t * x[n - 1];
        }
    }
}

/* ===========================================================================
 * CSF basis (kCSF = 1), npes = 1
 * ======================================================================== */

/* nonequiv_conf (src/csf.f90:454-536).  Configuration ic is equivalent to an
 * earlier jc when every entry of ic matches a distinct entry of jc with the
 * same l, 2j and occupation (first free match, in jc's entry order); then
 * ic's entries are reordered in place: for each entry position nj0 of jc, in
 * order, every entry of ic (scanned in order, no early exit) that matches jc's
 * entry nj0 is swapped into position nj0.  nc_neq[ic-1] receives the 1-based
 * class (the first equivalent configuration's class, else a new one).  The
 * reordering runs before Dinit, so Nq0 (per-entry occupation offsets) is
 * recomputed here (src/determinants.f90:84-103).  Returns the number of
 * classes. */
int orc_nonequiv(orc_ctx *c, int *nc_neq) {
    int *ind = xcalloc(c->Ns + 2, sizeof(int));
    int nneq = 0;
    for (int ic = 1; ic <= c->Nc; ic++) {
        const int i0 = c->Nc0[ic] + 1, i1 = c->Nc0[ic] + c->Nvc[ic];
        int found = 0;
        for (int jc = 1; jc < ic && !found; jc++) {
            const int j0 = c->Nc0[jc] + 1, j1 = c->Nc0[jc] + c->Nvc[jc];
            for (int njj = j0; njj <= j1; njj++) ind[c->Nip[njj]] = 0;
            int all = 1;
            for (int nii = i0; nii <= i1 && all; nii++) {
                const int ni = c->Nip[nii];
                int hit = 0;
                for (int njj = j0; njj <= j1; njj++) {
                    const int nj = c->Nip[njj];
                    if (ind[nj] == 1) continue;
                    if (c->Ll[ni] != c->Ll[nj]) continue;
                    if (c->Jj[ni] != c->Jj[nj]) continue;
                    if (c->Nq[nii] != c->Nq[njj]) continue;
                    ind[nj] = 1;
                    hit = 1;
                    break;
                }
                if (!hit) all = 0;
            }
            if (!all) continue;
            for (int nii = i0; nii <= i1; nii++) ind[c->Nip[nii]] = 0;
            const int nsj = j1 - j0 + 1, nsi = i1 - i0 + 1;
            for (int nj0 = 1; nj0 <= nsj; nj0++) {
                const int njj = nj0 + j0 - 1, nj = c->Nip[njj];
                for (int ni0 = 1; ni0 <= nsi; ni0++) {
                    const int nii = ni0 + i0 - 1, ni = c->Nip[nii];
                    if (ind[ni] == 1) continue;
                    if (c->Ll[ni] != c->Ll[nj]) continue;
                    if (c->Jj[ni] != c->Jj[nj]) continue;
                    if (c->Nq[nii] != c->Nq[njj]) continue;
                    if (nj0 != ni0) {
                        const int nii1 = nj0 + i0 - 1, ni1 = c->Nip[nii1];
                        int t = c->Nip[nii1]; c->Nip[nii1] = c->Nip[nii]; c->Nip[nii] = t;
                        t = c->Nq[nii1]; c->Nq[nii1] = c->Nq[nii]; c->Nq[nii] = t;
                        t = ind[ni1]; ind[ni1] = ind[ni]; ind[ni] = t;
                    }
                }
            }
            nc_neq[ic - 1] = nc_neq[jc - 1];
            found = 1;
        }
        if (!found) nc_neq[ic - 1] = ++nneq;
    }
    free(ind);
    /* Nq0 over the reordered entries (Dinit) */
    for (int ic = 1; ic <= c->Nc; ic++) {
        int n = 0;
        for (int ni = c->Nc0[ic] + 1; ni <= c->Nc0[ic] + c->Nvc[ic]; ni++) { c->Nq0[ni] = n; n += c->Nq[ni]; }
    }
    return nneq;
}

/* Number of determinants of configuration ic (1-based) with total 2M = m2,
 * by the Ndet odometer: N(Jm) - N(Jm + 1) is the number of J = Jm states of
 * the configuration. */
int orc_count_m(orc_ctx *c, int ic, int m2) {
    int *idet = xcalloc(c->Ne + 1, sizeof(int));
    int fin = 1, M, n = 0;
    ndet(c, ic, &fin, idet, &M);
    while (!fin) {
        if (M == m2) n++;
        ndet(c, ic, &fin, idet, &M);
    }
    free(idet);
    return n;
}

/* build_j2 (src/csf.f90:409-452): the dense J^2 block of configuration ic
 * (1-based) over its determinants, zz[(n-n1)*ndi + (k-n1)], both triangles
 * from the pairs k <= n, with csf.f90's inline F_J2 (nf = 0: Mj^2 plus the
 * shell terms; nf = 2: the Plj products; Rspq's outputs feed it as
 * Rspq_phase1/2's do). */
void orc_j2_block(const orc_ctx *c, int ic, double *zz) {
    const int Ne = c->Ne, ndi = c->Ndc[ic];
    int n1 = 0;
    for (int i = 1; i < ic; i++) n1 += c->Ndc[i];
    for (int i = 0; i < ndi * ndi; i++) zz[i] = 0.0;
    for (int n = 0; n < ndi; n++) {
        const int *d1 = c->iarr + (size_t)(n1 + n) * Ne;
        for (int k = 0; k <= n; k++) {
            int out[6];
            orc_rspq(Ne, d1, c->iarr + (size_t)(n1 + k) * Ne, out);
            if (k != n && out[0] != 2) continue;
            const double t = orc_fj2(c, d1, out[1], out[0], out[5], out[3], out[4], out[2]);
            if (t == 0.0) continue;
            zz[n * ndi + k] = t;
            zz[k * ndi + n] = t;
        }
    }
}

/* idif (src/csf.f90:1189-1227) on the configuration entries (Nip, Nq):
 * electrons that move between configurations ic and jc, 3 meaning more than
 * two. */
static int idif(const orc_ctx *c, int ic, int jc, int *iocc, int *jocc) {
    for (int n = 1; n <= c->Ns; n++) { iocc[n] = 0; jocc[n] = 0; }
    for (int nii = c->Nc0[ic] + 1; nii <= c->Nc0[ic] + c->Nvc[ic]; nii++) {
        if (nii == 0 || nii > c->Nsp) continue;
        iocc[c->Nip[nii]] = c->Nq[nii];
    }
    for (int njj = c->Nc0[jc] + 1; njj <= c->Nc0[jc] + c->Nvc[jc]; njj++) {
        if (njj == 0 || njj > c->Nsp) continue;
        jocc[c->Nip[njj]] = c->Nq[njj];
    }
    int i1 = 0, i2 = 0;
    for (int n = 1; n <= c->Ns; n++) {
        const int i = iocc[n] - jocc[n];
        if (i > 0) i1 += i;
        if (i < 0) i2 -= i;
        if (i1 > 2 || i2 > 2) return 3;
    }
    return i1;
}

/* compute_zzc_block (src/csf.f90:966-1006): zzc(ic, jc) = sum over the
 * determinants n of ic (in order) of buf_n(jc) * ccj(n, ic), buf_n(jc) = sum
 * over the determinants k of jc (in order, elements below 1e-20 in magnitude
 * skipped) of Hmatrix(n, k) * ccj(k, jc); E_k = Diag(k) with Kdsig. */
static void zzc_block(const orc_ctx *c, int n1, int ndi, int nci, const double *ci, int k1, int ndj, int ncj,
                      const double *cj, double *zz, double *buf) {
    const int Ne = c->Ne;
    for (int i = 0; i < nci * ncj; i++) zz[i] = 0.0;
    for (int n = 0; n < ndi; n++) {
        const int *d1 = c->iarr + (size_t)(n1 + n) * Ne;
        for (int jc = 0; jc < ncj; jc++) buf[jc] = 0.0;
        for (int k = 0; k < ndj; k++) {
            int out[6];
            orc_rspq(Ne, d1, c->iarr + (size_t)(k1 + k) * Ne, out);
            if (out[0] > 2) continue;  /* Hmatrix gives 0 */
            if (c->Kdsig != 0) ((orc_ctx *)c)->E_k = c->DiagS[k1 + k + 1];
            const double hij = hmltn_core(c, d1, out[1], out[0], out[5], out[3], out[4], out[2], 0);
            if (fabs(hij) < 1e-20) continue;
            for (int jc = 0; jc < ncj; jc++) buf[jc] = buf[jc] + hij * cj[k + (size_t)jc * ndj];
        }
        for (int ic = 0; ic < nci; ic++)
            for (int jc = 0; jc < ncj; jc++) zz[ic * ncj + jc] = zz[ic * ncj + jc] + buf[jc] * ci[n + (size_t)ic * ndi];
    }
}

/* formh_sym, npes = 1 (src/csf.f90:538-964), with RedistributeHamCSR's
 * stable counting sort by row (src/matrix_io.f90:445-489): H in the CSF
 * basis, lower triangle, every (iconf >= jconf, idif <= 2) block stored in
 * full (jc <= ic on the diagonal block), zeros included.  Row n = ic +
 * mdcs(iconf) holds jconf ascending, jc ascending.
 *   nc_neq[Nc] (1-based classes), ndcs[nneq], place[nneq] (0-based offsets
 *   of each class's coefficients in ccj, column-major ndi x ndcs).
 * Only the rows of configurations ic_lo .. ic_hi (1-based; 0 for all) are
 * formed.  Count mode (col == NULL): counts[ncsf] gets the row lengths.  Fill
 * mode: rowptr[ncsf+1] from 0, col (1-based) and val filled; *numzero =
 * stored zeros, *hmin = minval of the stored values (src/csf.f90:934),
 * *npairs = configuration pairs evaluated, *ncand = determinant pairs in
 * them.  Returns the number of stored elements. */
int64_t orc_formh_sym(const orc_ctx *c, const int *nc_neq, const int *ndcs, const int64_t *place,
                      const double *ccj, int ic_lo, int ic_hi, int64_t *counts, const int64_t *rowptr, int *col,
                      double *val, int64_t *numzero, double *hmin, int64_t *npairs, int64_t *ncand) {
    const int Nc = c->Nc;
    int *mdcs = xcalloc(Nc + 2, sizeof(int)), *md = xcalloc(Nc + 2, sizeof(int));
    int ncsf = 0, maxc = 0;
    for (int ic = 1; ic <= Nc; ic++) {
        mdcs[ic] = ncsf;
        const int q = nc_neq[ic - 1] - 1;
        ncsf += ndcs[q];
        if (ndcs[q] > maxc) maxc = ndcs[q];
        md[ic + 1] = md[ic] + c->Ndc[ic];
    }
    int *iocc = xcalloc(c->Ns + 1, sizeof(int)), *jocc = xcalloc(c->Ns + 1, sizeof(int));
    int64_t *cur = xcalloc(ncsf + 1, sizeof(int64_t));
    double *zz = xcalloc((size_t)maxc * maxc + 1, sizeof(double)), *buf = xcalloc(maxc + 1, sizeof(double));
    if (counts) for (int r = 0; r < ncsf; r++) counts[r] = 0;
    int64_t ih8 = 0, nz = 0, np = 0, nk = 0;
    double mn = 0.0;
    if (ic_lo < 1) ic_lo = 1;
    if (ic_hi < 1 || ic_hi > Nc) ic_hi = Nc;
    for (int ic = ic_lo; ic <= ic_hi; ic++) {
        const int qi = nc_neq[ic - 1] - 1, nci = ndcs[qi];
        if (nci == 0) continue;
        const int ndi = c->Ndc[ic];
        for (int jc = 1; jc <= ic; jc++) {
            const int qj = nc_neq[jc - 1] - 1, ncj = ndcs[qj];
            if (ncj == 0) continue;
            const int ndj = c->Ndc[jc];
            if (idif(c, ic, jc, iocc, jocc) > 2) continue;
            np++;
            nk += (int64_t)ndi * ndj;
            if (col) zzc_block(c, md[ic], ndi, nci, ccj + place[qi], md[jc], ndj, ncj, ccj + place[qj], zz, buf);
            for (int a = 0; a < nci; a++) {
                const int jc2 = (ic == jc) ? a + 1 : ncj;
                const int row = a + mdcs[ic];
                for (int b = 0; b < jc2; b++) {
                    if (col) {
                        const int64_t pos = rowptr[row] + cur[row]++;
                        const double h = zz[a * ncj + b];
                        col[pos] = b + mdcs[jc] + 1;
                        val[pos] = h;
                        if (h == 0.0) nz++;
                        if (ih8 == 0 || h < mn) mn = h;
                    } else if (counts) {
                        counts[row]++;
                    }
                    ih8++;
                }
            }
        }
    }
    free(mdcs); free(md); free(iocc); free(jocc); free(cur); free(zz); free(buf);
    if (numzero) *numzero = nz;
    if (hmin) *hmin = mn;
    if (npairs) *npairs = np;
    if (ncand) *ncand = nk;
    return ih8;
}
