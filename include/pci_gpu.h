/*
 * pci_gpu.h -- C ABI of the MI355X pconf hot path (H build + Davidson).
 *
 * The reference (ud-pci/pCI) has no C ABI: its hot path is three Fortran
 * module procedures that share state through module globals
 * (src/conf_variables.F90:9-73):
 *
 *   FormH(npes, mype)                    src/pconf.F90:1160-1651
 *   RedistributeHamCSR(npes, mype, ...)  src/matrix_io.f90:260-502
 *   FormJ / RedistributeJsqCSR / J_av    src/formj2.f90:69-387, src/matrix_io.f90:504-687
 *   Diag4(mype, npes)                    src/pconf.F90:1993-2264 (+ src/davidson.f90)
 *
 * Each entry point below replaces one of them (or the broadcast of its
 * inputs in InitFormH, src/pconf.F90:1049-1158) for a single GPU.  Arrays are
 * plain host pointers owned by the caller; device memory belongs to the
 * context.  Index conventions are the reference's: positions, shells and
 * determinants are 1-based, Iarr is Fortran Iarr(Ne, Nd) (determinant-major).
 * Every call returns 0 on success or a negative PCI_E* code; the text of the
 * last error of the calling thread is returned by pci_gpu_last_error().
 * One host thread and one HIP stream per context.
 */
#ifndef PCI_GPU_H
#define PCI_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PCI_OK            0
#define PCI_E_ARG        -1   /* invalid argument or call order            */
#define PCI_E_HIP        -2   /* HIP runtime / library error               */
#define PCI_E_NOMEM      -3   /* device allocation failed                  */
#define PCI_E_INTEGRAL   -4   /* "Gint: integral is absent" (integrals.f90:256-261) */
#define PCI_E_ORTHO      -5   /* "Fail of orthogonalization" (davidson.f90:717-723) */
#define PCI_E_UNSUPPORTED -6  /* input outside what this build handles     */

typedef struct pci_ctx pci_ctx;

/* Context lifetime.  device: HIP device ordinal. */
int pci_gpu_create(int device, pci_ctx **ctx);
int pci_gpu_destroy(pci_ctx *ctx);
const char *pci_gpu_last_error(void);

/* Orbital basis, spin-orbital table and determinant list.
 * Replaces the InitFormH broadcast of Nn/Ll/Kk/Jj, Jz/Nh, Iarr and Ndc
 * (src/pconf.F90:1128-1142) that FormH reads.
 *   ns, nso          : number of shells, core shells
 *   ll, jj, kk       : per shell (length ns): l, 2j, kappa
 *   npos, jz, nh     : per spin-orbital position (length npos): 2m and shell
 *   ne, nd, iarr     : Iarr(Ne, Nd), sorted 1-based positions per determinant
 *   nc, ndc          : determinants per relativistic configuration, Ndc(Nc)
 *   mj               : 2|Jm| (Mj of src/pconf.F90:429)
 *   gj               : Gj, coefficient of the J^2 term in Hmltn            */
int pci_gpu_set_space(pci_ctx *ctx, int32_t ns, int32_t nso, const int32_t *ll,
                      const int32_t *jj, const int32_t *kk, int32_t npos,
                      const int32_t *jz, const int32_t *nh, int32_t ne, int32_t nd,
                      const int32_t *iarr, int32_t nc, const int32_t *ndc, int32_t mj,
                      double gj);

/* Integral tables as pconf holds them after Rint and Input
 * (CONF.INT: src/integrals.f90:35-83; CONF.GNT: src/pconf.F90:382-406),
 * replacing BuildHintLUT/BuildGintHash/BuildGauntLUT (integrals.f90:604-715).
 *   nx                : index radix Nx (= sqrt of CONF.INT record 7 nrd)
 *   nhint,rint1,iint1 : one-electron integrals and their indices
 *   ngint,kbrt,rint2  : two-electron integrals, FP64, Rint2(1|2, Ngint)
 *   iint2, iint3      : their (iac, ibd) keys
 *   nlx,ngaunt,in,gnt : Gaunt table                                         */
int pci_gpu_set_integrals(pci_ctx *ctx, int32_t nx, int32_t nhint, const double *rint1,
                          const int32_t *iint1, int64_t ngint, int32_t kbrt,
                          const double *rint2, const int32_t *iint2, const int32_t *iint3,
                          int32_t nlx, int32_t ngaunt, const int32_t *in, const double *gnt);

/* MBPT one-electron effective integrals, Ksig = 1 or 2 of a Kl = 2 run: the
 * SGC.CON tables RintS reads (src/pconf.F90:715-762) -- nhints pairs with
 * Iint1S = Nx*(na-Nso-1)+(nb-Nso), Rsig, Dsig, Esig -- NmaxS, LmaxS, the
 * orbital energies Eps(1..Ns) of CONF.DAT (src/pconf.F90:663-667) and, for
 * Kdsig != 0, FormD's Diag(1..Nd) = E_0 - sum of Eps (src/pconf.F90:852-880;
 * NULL otherwise).  pci_gpu_build_h then adds HintS (src/integrals.f90:
 * 321-369, Kexn = 1) to every Hint with E_k = Diag of the column determinant
 * (src/pconf.F90:1275).  ksig = 0 switches the correction off; ksig = 2
 * also needs pci_gpu_set_sigma2.  Call after pci_gpu_set_space and
 * pci_gpu_set_integrals.                                                  */
int pci_gpu_set_sigma(pci_ctx *ctx, int32_t ksig, int32_t kdsig, int32_t nmaxs, int32_t lmaxs,
                      int32_t nhints, const int32_t *iint1s, const double *rsig, const double *dsig,
                      const double *esig, const double *eps, const double *diag);

/* MBPT with screening (Ksig = 2): the second half of RintS
 * (src/pconf.F90:764-840) -- SGC.CON's ten screening coefficients scr
 * (multipoles k = 0..9) and SCRC.CON: Ksym, Kmax, Nsum (4 Nav, or 4 NmaxS
 * when Nsx2 != 0) and the two-electron sigma integrals (ngints radial values
 * rsig2/dsig2/esig2 with their Iint2S/Iint3S keys, 1-based).  Call after
 * pci_gpu_set_sigma with ksig = 2 and before pci_gpu_build_h; Gint then
 * screens R^k and adds GintS (src/integrals.f90:262-271, 371-493).        */
int pci_gpu_set_sigma2(pci_ctx *ctx, int32_t ksym, int32_t kmax, int32_t nsum, const double *scr,
                       int64_t ngints, const int32_t *iint2s, const int32_t *iint3s, const double *rsig2,
                       const double *dsig2, const double *esig2);

/* Dense symmetric eigensolver of the fast mode's Init4 block, exposed for
 * callers and tests: the DSYEVD of DiagInitApprox (src/pconf.F90:1876-1889)
 * in FP32.  a: n x n column-major (both triangles), overwritten by the
 * eigenvectors; w: the n eigenvalues, ascending.  *path (may be NULL) is 1
 * when the one-launch cooperative tridiagonalisation ran, 0 for rocSOLVER's
 * ssyevd (matrices above 3072 or beyond the GPU's LDS, or
 * PCI_EIGH_ROCSOLVER set), 2 when the cooperative kernel gave up at a grid
 * barrier whose wait ran out and ssyevd solved the untouched input instead.
 * No counterpart in the reference's interface.                              */
int pci_gpu_eigh32(pci_ctx *ctx, int32_t n, float *a, float *w, int32_t *path);

/* FormH + RedistributeHamCSR for one rank (src/pconf.F90:1226-1356,
 * 1597-1627; src/matrix_io.f90:260-502).  Builds H on the device and keeps
 * it resident.  Outputs: numh = stored lower-triangle elements (diagonal
 * included, exact zeros dropped as pconf does at npes=1), numzero = dropped
 * pairs, hmin = Hamil%minval.  flags: 0 = the CSR (and the SpMM copies the
 * autotuner picks), or PCI_BUILD_BLOCKS.                                    */
#define PCI_BUILD_BLOCKS 1  /* build H straight into the residual blocks the
                               SpMM streams (FP64, and FP32 with the shifted
                               diagonal: 12 B per block value), no CSR: for
                               spaces whose CSR and blocks do not fit
                               together.  Same elements and statistics; the
                               CSR read-back (pci_gpu_get_csr, _rows) is
                               assembled from the blocks; SpMM block widths
                               up to 24; 2 to 8 valence electrons.          */
int pci_gpu_build_h(pci_ctx *ctx, int32_t flags, int64_t *numh, int64_t *numzero,
                    double *hmin);

/* Statistics of the last pci_gpu_build_h: ncand = determinant pairs whose
 * element was evaluated (both triangles, exact zeros included), nnz_full =
 * elements stored in both triangles, build_ms = wall time of the build.
 * pconf prints the like as FormH's TIMING and progress lines
 * (src/pconf.F90:1291-1315); no counterpart in its module interface.      */
int pci_gpu_build_info(pci_ctx *ctx, int64_t *ncand, int64_t *nnz_full, double *build_ms);

/* Copy the lower-triangle CSR out in pCONF.HIJ order
 * (HamilCSR_rowptr(0:Nd) 0-based int64, Hamil%col 1-based int32,
 * Hamil%val FP64; src/matrix_io.f90:689-799).  Buffers sized nd+1 / numh. */
int pci_gpu_get_csr(pci_ctx *ctx, int64_t *rowptr, int32_t *col, double *val);

/* Rows row1 .. row1+nrows-1 (1-based) of the lower-triangle CSR, as one
 * rank of WriteMatrixCSR writes its own row range (src/matrix_io.f90:
 * 689-799): rowptr[nrows+1] starts at 0; col (1-based) and val receive
 * rowptr[nrows] elements when given (cap = their length; pass col = val =
 * NULL to read the row pointers and size the buffers first).              */
int pci_gpu_get_csr_rows(pci_ctx *ctx, int32_t row1, int32_t nrows, int64_t *rowptr, int32_t *col,
                         double *val, int64_t cap);

/* Load a prebuilt lower-triangle CSR instead of building it: the Kl=1 path
 * (ReadMatrixCSR, src/matrix_io.f90:801-939; src/pconf.F90:201-215). */
int pci_gpu_set_csr(pci_ctx *ctx, int32_t nd, int64_t numh, const int64_t *rowptr,
                    const int32_t *col, const double *val);

/* FormJ + RedistributeJsqCSR (src/formj2.f90:120-184): J^2 on the device. */
int pci_gpu_build_j(pci_ctx *ctx, int64_t *numj);
int pci_gpu_get_jcsr(pci_ctx *ctx, int64_t *rowptr, int32_t *col, double *val);

/* CSF basis, kCSF = 1 (src/pconf.F90:146-149, 211-214, 247-255).
 *
 * pci_gpu_csf_basis replaces jbasis (src/csf.f90:17-407): for each class of
 * equivalent configurations (nonequiv_conf, src/csf.f90:454-536, run by the
 * caller before the determinants are listed, so that the configurations of
 * a class list their determinants in corresponding order) the J^2 block of
 * its first configuration (build_j2) is diagonalised and the eigenvectors
 * with J = |Jm| are kept resident as the class's CSFs.
 *   nconf_neq, nc_neq[nc] : number of classes, 1-based class of each
 *                           configuration (nc_neq of nonequiv_conf)
 *   ndcs[nconf_neq]       : out, CSFs of each class (NULL allowed)
 *   ncsf, nccj            : out, CSFs of the space (pconf's Ncsf) and stored
 *                           coefficients (nccj)
 * An eigenvalue of J^2 that is not J(J+1) to 1e-7 for a half-integer J is
 * "Value of j in jbasis is wrong" (PCI_E_ARG), as jbasis aborts.  The
 * eigenvectors of a degenerate J are any orthonormal basis of their space
 * (rocSOLVER's, not LAPACK's), so they can differ from pconf's by a rotation
 * inside each class; the eigenvalues of H in the basis do not.            */
int pci_gpu_csf_basis(pci_ctx *ctx, int32_t nconf_neq, const int32_t *nc_neq, int32_t *ndcs, int32_t *ncsf,
                      int64_t *nccj);

/* The resident coefficients, ccj_module's layout (src/csf.f90:318-350):
 * class after class, ndcs columns of the class's Ndc values each
 * (column-major).  ccj: nccj values.  No counterpart in the reference's
 * interface (pconf keeps them in module memory).                          */
int pci_gpu_csf_get_basis(pci_ctx *ctx, double *ccj);

/* Install CSF coefficients computed elsewhere -- pconf's own ccj_module
 * from its jbasis, say -- in place of pci_gpu_csf_basis's: the classes
 * nc_neq[nc], CSFs per class ndcs[nconf_neq] and the coefficients in
 * ccj_module's layout (nccj = sum of ndcs x the class's Ndc values).  On the
 * same coefficients pci_gpu_build_h_csf gives pconf's formh_sym values bit
 * for bit.  No counterpart in the reference's interface.                    */
int pci_gpu_csf_set_basis(pci_ctx *ctx, int32_t nconf_neq, const int32_t *nc_neq, const int32_t *ndcs,
                          const double *ccj);

/* formh_sym + RedistributeHamCSR (src/csf.f90:538-964): H in the CSF basis
 * (Ncsf x Ncsf) becomes the context's resident H, so pci_gpu_davidson then
 * solves in that basis and pci_gpu_get_csr returns its lower triangle.
 * Every block of configurations that differ by at most two electrons is
 * stored in full, zeros included (numzero counts them), as pconf stores
 * them; hmin = the least stored value (Hamil%minval, src/csf.f90:934).
 * Hmatrix has no Gj term in this basis.  Needs pci_gpu_csf_basis and the
 * integrals.                                                               */
int pci_gpu_build_h_csf(pci_ctx *ctx, int64_t *numh, int64_t *numzero, double *hmin);

/* unsym (src/csf.f90:1088-1134): determinant-basis vectors y[nvec*nd] from
 * CSF-basis vectors x[nvec*ncsf] (both vector-major).                       */
int pci_gpu_csf_unsym(pci_ctx *ctx, int32_t nvec, const double *x, double *y);

/* Target J and the J terms for the J selection of Kv = 3: XJ_av ("J" of
 * CONF.INP, src/conf_init.f90:74-75) and Jt(1:Njd), the doubled J values of
 * the space in ascending order as Jterm computes them
 * (src/determinants.f90:185-348); used with PCI_DVD_JSEL.                   */
int pci_gpu_set_jterms(pci_ctx *ctx, double xj_av, int32_t njd, const int32_t *jt);

/* Diag4 (src/pconf.F90:1993-2264): block Davidson for the lowest nlv roots.
 * The initial approximation is the dense eigenproblem of the leading
 * nd0 x nd0 block (Init4 + DiagInitApprox, davidson.f90:20-73,
 * pconf.F90:1822-1991), solved on the device.                               */
#define PCI_DVD_FP32       1  /* FP32-stored H, FP32 SpMM, FP64 refinement      */
#define PCI_DVD_LEARNED    2  /* MLP gain on the residual (on the coupled one
                                 with PCI_DVD_COUPLED), rows > nd0; a step that
                                 does not reduce the residual is undone and the
                                 diagonal correction is used from then on    */
                          /* 4: not used                                   */
#define PCI_DVD_TRACE      8  /* record per-iteration residuals in the stats   */
#define PCI_DVD_COUPLED   16  /* rows > nd0: (r + H[tail,head] t_head)/(theta-D) */
#define PCI_DVD_JSEL      32  /* Kv = 3 (K_prj = 1): J filter in FormB0, AvgDiag and
                                 Prj_J (needs pci_gpu_build_j, pci_gpu_set_jterms) */
#define PCI_DVD_NOADMIX   64  /* stop on the residual alone: no admixture test
                                 (ax <= 1e-6, pconf.F90:2162-2180); the solve
                                 runs until max ||r|| < crt4 or n_it          */

typedef struct {
    int32_t nlv;     /* Nlv                                  */
    int32_t nd0;     /* Nd0 (calcNd0, determinants.f90:22-45) */
    double crt4;     /* Crt4: residual norm for convergence  */
    int32_t n_it;    /* N_it: iteration cap                  */
    int32_t flags;   /* PCI_DVD_* bits                       */
    int32_t refine_every; /* FP32 mode: FP64 H.x refresh period (0 = auto) */
    int32_t reserved;
} pci_dvd_params;

typedef struct {
    int32_t iterations;      /* Davidson iterations performed            */
    int32_t converged;       /* 1: max FP64 residual <= crt4; 2: stopped by the
                                admixture test (pconf.F90:2162-2180); 0: N_it reached */
    int32_t spmm_fp32;       /* FP32 SpMM launches                        */
    int32_t spmm_fp64;       /* FP64 SpMM launches                        */
    int32_t learned_used;    /* iterations whose MLP correction was kept   */
    int32_t learned_rejected;/* learned iterations undone: the rejected one and
                                the provisional ones before it              */
    double max_residual;     /* final FP64 residual norm (max over roots) */
    double t_total_ms;       /* device time of the whole solve            */
    double t_spmm_ms;        /* summed SpMM kernel time (HIP events around the
                                SpMM launches, conversions excluded)     */
    double t_init_ms;        /* Init4 + initial eigensolve                */
    double spmm_bytes;       /* algorithmic bytes of all SpMM launches    */
    double t_spmm32_ms;      /* the FP32 launches' share of t_spmm_ms     */
    double resid_hist[512];  /* per-iteration max residual (PCI_DVD_TRACE); with
                                PCI_DVD_LEARNED, the residual a learned step
                                reached before it was undone              */
} pci_dvd_stats;

/* Runs Diag4.  Outputs (caller-allocated): tk[nlv] = -(E + Hmin) as written
 * to CONF.XIJ, tj[nlv] = <J> from J_av (needs pci_gpu_build_j, else 0),
 * vecs[nlv*nd] eigenvectors (vector-major).  stats may be NULL. */
int pci_gpu_davidson(pci_ctx *ctx, const pci_dvd_params *p, double *tk, double *tj,
                     double *vecs, pci_dvd_stats *stats);

/* Checkpoint hook of a solve: called on the calling thread with the current
 * approximations -- tk[nlv] (as CONF.XIJ holds them), tj[nlv] (J_av, 0
 * without J^2), vecs[nlv*nd] (vector-major) -- after FormB0 (it = 0) and
 * after every `every`-th iteration.  The buffers are valid during the call. */
typedef void (*pci_dvd_checkpoint_fn)(void *user, int32_t it, int32_t nlv, int32_t nd, const double *tk,
                                      const double *tj, const double *vecs);

/* Diag4 with pconf's restart and periodic output:
 *   x0[nlv*nd] (vector-major, may be NULL), tk0[nlv] (may be NULL): start
 *     from these vectors instead of the Nd0 block's eigenvectors -- FormB0
 *     with |Kl4| = 2, which reads them from CONF.XIJ (src/davidson.f90:
 *     136-154; p->nlv = the number read); Init4 still runs (its eigenpairs
 *     shape the correction);
 *   every > 0 with cb: the checkpoint above, as FormB writes CONF.XIJ every
 *     KXIJ iterations (src/pconf.F90:2186-2211, src/davidson.f90:219-263).
 * Otherwise as pci_gpu_davidson.                                            */
int pci_gpu_davidson_ex(pci_ctx *ctx, const pci_dvd_params *p, const double *tk0, const double *x0, int32_t every,
                        pci_dvd_checkpoint_fn cb, void *user, double *tk, double *tj, double *vecs,
                        pci_dvd_stats *stats);

/* Batched Diag4 over independent CI blocks (the J / parity blocks of one
 * atom, each in its own context): every context's solve runs at once, one
 * host thread per context on the context's own stream, so the blocks share
 * the GPU.  p, tk, tj, vecs, stats: one entry per block (tk, tj, vecs,
 * stats may be NULL as a whole).  All contexts on one device, none twice.
 * The part of the head eigensolve (Init4) that needs the whole GPU runs
 * one block at a time while the other blocks wait with idle streams; the
 * rest of the eigensolve and the iterations of all blocks overlap.
 * Each block's result is bit-identical to its own pci_gpu_davidson.  A
 * block that fails does not stop the others; the call returns the first
 * failing block's error.  No counterpart in the reference (pconf solves one
 * block per run; the J / parity blocks of an atom are separate runs).      */
int pci_gpu_davidson_batch(int32_t nblocks, pci_ctx *const *ctx, const pci_dvd_params *p,
                           double *const *tk, double *const *tj, double *const *vecs, pci_dvd_stats *stats);

/* One application Y = (H - Hmin I) X of the resident matrix to a block of
 * nvec vectors (Mxmpy, davidson.f90:572-647), for measurement and tests.
 * x, y: host arrays nd*nvec, row-major (x[i*nvec + v]).  precision: 32 or 64.
 * Supported widths nvec: 1..32, 40, 48 and 64 (others: PCI_E_ARG).
 * reps > 1 repeats the launch on resident data; *ms gets the mean kernel
 * time from HIP events on the context stream, *bytes the algorithmic bytes
 * of the kernel in use: the CSR stream (value + 4-byte column per element,
 * row pointers) for the CSR and sliced kernels, the block values, list
 * gathers and list-slot sums for the residual-block kernel (no column
 * indices); X read and Y written once either way.                          */
int pci_gpu_spmm(pci_ctx *ctx, int32_t nvec, int32_t precision, const double *x, double *y,
                 int32_t reps, double *ms, double *bytes);

/* Allocation-free form of pci_gpu_spmm, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.
 * x, y: device arrays, nd x nvec row-major with leading dimensions ldx, ldy
 * (FP32 for precision 32, FP64 for 64), rows 16-byte aligned and
 * ldx, ldy >= nvec rounded up to 4 (FP32) or 2 (FP64) values; y = (H - Hmin I) x.
 * Runs the kernel the autotuner already chose for (nvec, precision) on its
 * copy of H, with the product's partial sums in a 256 MiB device array of
 * the library (each GPU holds it until the process ends): calls no
 * hipMalloc/hipFree, builds no copy, and returns PCI_E_ARG if the kernel is
 * not chosen or its copy not built yet (call pci_gpu_spmm, or a Davidson
 * solve, with the same nvec and precision first), if precision 32 is asked
 * before the FP32 values exist, or if the product needs more than 256 MiB
 * of partial sums.  Synchronous.  Calls on one GPU share the array: calls
 * that overlap on one GPU (from several threads) race.                      */
int pci_gpu_spmm_noalloc(pci_ctx *ctx, int32_t nvec, int32_t precision, const void *x, int32_t ldx, void *y,
                         int32_t ldy);

/* Name of the SpMM kernel the autotuner chose for this block width and
 * precision on the resident matrix ("" before the first use).  The first
 * SpMM of each (precision, nvec) on a new matrix times the candidate kernel
 * shapes on that call's operands and keeps the fastest.  No counterpart in
 * the reference (Mxmpy has one loop shape).                                */
int pci_gpu_spmm_variant(pci_ctx *ctx, int32_t nvec, int32_t precision, char *name, int32_t len);
/* Pin the kernel shape (0 .. *count-1) used for (nvec, precision) instead of
 * the autotuner's choice, e.g. for reproducible timings; variant < 0 only
 * returns the number of shapes in *count.  No counterpart in the reference. */
int pci_gpu_spmm_set_variant(pci_ctx *ctx, int32_t nvec, int32_t precision, int32_t variant, int32_t *count);

/* Learned correction (MLP preconditioner).  No counterpart in the
 * reference: it stands in for the diagonal part of the Dvdsn correction
 * (src/davidson.f90:482-488) and is only used with PCI_DVD_LEARNED.
 * Network: x (8 features + 16-wide embedding) -> 128 -> 128 -> 1, BF16
 * weights with FP32 accumulation; parameters are a flat FP32 array of
 * pci_gpu_mlp_nparams values (W1 128x32, b1, W2 128x128, b2, w3, b3,
 * E 64x16).  Samples: x[n*8] features, bucket[n], r[n] residual
 * components, target[n] wanted corrections; the loss is the mean of
 * (r * g - target)^2 with g the network's gain.                            */
int pci_gpu_mlp_nparams(int64_t *n);
int pci_gpu_mlp_init(pci_ctx *ctx, uint64_t seed);
int pci_gpu_mlp_set_weights(pci_ctx *ctx, const float *w, int64_t n);
int pci_gpu_mlp_get_weights(pci_ctx *ctx, float *w, int64_t n);
/* Training pairs from a Davidson run against reference eigenvectors xref
 * (nlv x nd): every stride-th determinant >= nd0, every root, every
 * iteration, up to cap samples; *n receives the count.                     */
int pci_gpu_mlp_collect(pci_ctx *ctx, const pci_dvd_params *p, const double *xref, int32_t stride,
                        int64_t cap, float *x, int32_t *bucket, float *r, float *target, int64_t *n);
/* Adam training for `epochs` passes in batches of `batch` samples (BF16
 * forward/backward kernels); loss_hist[epochs] receives the epoch losses. */
int pci_gpu_mlp_train(pci_ctx *ctx, const float *x, const int32_t *bucket, const float *r,
                      const float *target, int64_t n, int32_t epochs, int32_t batch, float lr,
                      float *loss_hist);
/* Loss and gradient over n samples (grad: pci_gpu_mlp_nparams values).    */
int pci_gpu_mlp_grad(pci_ctx *ctx, const float *x, const int32_t *bucket, const float *r,
                     const float *target, int64_t n, float *grad, float *loss);
/* Corrections t = r * g for n samples.                                     */
int pci_gpu_mlp_forward(pci_ctx *ctx, const float *x, const int32_t *bucket, const float *r,
                        int64_t n, float *t);

/* Device memory held by the context, in bytes. */
int pci_gpu_mem_usage(pci_ctx *ctx, int64_t *bytes);

#ifdef __cplusplus
}
#endif
#endif /* PCI_GPU_H */
