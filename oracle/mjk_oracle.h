/*
 * CPU FP64 restatement of OpenMolcas's Cholesky J/K Fock build.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity reference for the HIP
 * kernels in openmolcas_amd/csrc and the CPU baseline timed by bench.py.  The
 * product path never links, loads or calls it.
 *
 * Conventions (nSym = 1, JSYM = 1):
 *   - LT packed storage, 0-based: element (a,b), a >= b, at a*(a+1)/2 + b
 *     (src/misc_util/index_functions.F90:138-146, 1-based iTri(i,j) = i(i-1)/2+j).
 *   - Densities passed for Coulomb are "folded": off-diagonal elements doubled
 *     (src/misc_util/fold.F90:14); Fock matrices are not folded.
 *   - Orbital matrices are the reference's Porb(d)%SB(1)%A2(nOcc,nBas): the
 *     orbital index runs fastest, element (k,a) at k + a*nOcc.
 *   - A Cholesky vector J of reduced set r is stored on the nrs[r] pairs of
 *     that set; rs2lt maps each reduced-set position to its 0-based LT index
 *     (the composition IndRed(.,3) -> iRS2F -> iTri of
 *     src/fock_util/swap_full2rs.F90:37-46).
 */
#ifndef MJK_ORACLE_H
#define MJK_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Vector store as the consumer sees it after Cho_X_Init
 * (src/cholesky_util/cho_x_init.F90:55): reduced sets, their index maps and
 * the vectors ordered by reduced set. */
typedef struct {
  int32_t n;             /* nBas */
  int32_t nred;          /* number of reduced sets */
  const int64_t *nrs;    /* [nred] nDimRS(1,r) */
  const int64_t *rs_off; /* [nred] offset of set r's map inside rs2lt */
  const int64_t *rs2lt;  /* concatenated maps, 0-based LT indices */
  int64_t nvec;          /* NumCho(1) */
  const int32_t *vred;   /* [nvec] reduced set of each vector (InfVec(J,2)) */
  const int64_t *voff;   /* [nvec] offset of vector J inside lvec */
  const double *lvec;    /* concatenated vectors, nrs[vred[J]] values each */
} ora_chol;

void ora_set_threads(int32_t n);
/* Timing only: a Fortran BLAS dgemv (32-bit integers) for ora_dgemv and
 * ora_dgemm_tri_tn, or NULL for the plain loops (the default). */
void ora_set_dgemv(void *fn);
int32_t ora_get_threads(void);

int64_t ora_itri(int64_t i, int64_t j); /* 1-based, either order */
void ora_fold(int32_t n, const double *a_sq, double *b_lt);
void ora_unfold(int32_t n, const double *a_lt, double *b_sq);

void ora_swap_full2rs(int64_t nrs, const int64_t *rs2lt, int32_t nmat,
                      int64_t ldlt, const double *xlt, double *xrs);
void ora_swap_rs2full(int64_t nrs, const int64_t *rs2lt, int32_t nmat,
                      int64_t ldlt, const double *xrs, double *xlt,
                      int32_t add);

void ora_dgemv(char trans, int64_t m, int64_t n, double alpha,
               const double *a, int64_t lda, const double *x, double beta,
               double *y);
void ora_dgemm_tri_tn(int64_t n, int64_t k, double alpha, const double *a,
                      int64_t lda, const double *b, int64_t ldb, double beta,
                      double *c_lt);
void ora_dgemm_tri_nt(int64_t n, int64_t k, double alpha, const double *a,
                      int64_t lda, const double *b, int64_t ldb, double beta,
                      double *c_lt);

void ora_cho_vtra2(int32_t n, int64_t nrs, const int64_t *rs2lt,
                   int64_t jnum, const double *lrs, int32_t norb,
                   const double *mo, double *chot);
void ora_cho_vtra0(int32_t n, int64_t nrs, const int64_t *rs2lt,
                   int64_t jnum, const double *lrs, int32_t norb,
                   const double *mo, double *chot);

int32_t ora_cho_fscf(const ora_chol *ch, int32_t nden, double *flt,
                     const int32_t *nforb, const int32_t *niorb,
                     const double *const *porb, const double *dlt,
                     double exfac, int64_t nvec_batch);

int32_t ora_cho_fock_dft_red(const ora_chol *ch, const double *dlt, double *flt,
                             int64_t nvec_batch);

int32_t ora_cho_lk_scf(const ora_chol *ch, int32_t nden, double *flt, double *klt,
                       const int32_t *nforb, const int32_t *niorb,
                       const double *const *porb, const double *plt, double factxi,
                       int32_t nshell, const int32_t *nbassh, int64_t nvec_batch);

int32_t ora_cho_fmcscf(const ora_chol *ch, double *flt, int32_t nforb,
                       int32_t niorb, int32_t naorb, double factxi,
                       const double *dlt, int32_t do_active,
                       const double *porb1, int32_t nchm,
                       const double *porb2, const double *porb3,
                       double *w_pwxy, const double *cmo, int32_t nfro,
                       int32_t norb, double exfac, int64_t nvec_batch);

int32_t ora_cho_lk_casscf(const ora_chol *ch, const double *dlt, double *flt,
                          const double *msq, double *w_pwxy, double factxi,
                          int32_t nfiorb, int32_t naorb, int32_t nchm,
                          const double *ash1, const double *ash2, int32_t do_active,
                          const double *cmo, int32_t nfro, int32_t norb, double exfac,
                          int64_t nvec_batch);

/* LK screening state of CHO_LK_SCF / CHO_LK_CASSCF that the reference reads
 * from module globals: the shells (nShell, nBasSh; the AOs of a shell are
 * contiguous, shells in order), the diagonal CHO_IODIAG reads (LT storage,
 * zero outside the first reduced set), ThrCom, MaxRed and the
 * Fock_util_global switches Estimate / Update (src/fock_util/
 * fock_util_global.F90:56-59).  nscreen is set from the call's argument. */
typedef struct {
  int32_t nshell;
  const int32_t *nbassh;
  const double *diag;
  double thrcom;
  int32_t maxred;
  int32_t estimate, update;
  int32_t nscreen;
} ora_lk;

void ora_lk_thresholds(double thrcom, double dfmat, double dmpk, int32_t nden,
                       const int32_t *nocc, int64_t numvt, int32_t estimate,
                       int32_t maxred, double *tau, double *thrv);

int32_t ora_lk_exchange(const ora_chol *ch, int32_t nden, const int32_t *nocc,
                        const double *const *csq, const double *fact, const double *tau,
                        const double *thrv, int32_t cas_cut, const ora_lk *lk,
                        int64_t nvec_batch, double *klt, int64_t *stats);

int32_t ora_cho_lk_scf_scr(const ora_chol *ch, int32_t nden, double *flt, double *klt,
                           const int32_t *nforb, const int32_t *niorb,
                           const double *const *porb, const double *plt, double factxi,
                           int32_t nscreen, double dmpk, double dfmat, const ora_lk *lk,
                           int64_t nvec_batch, int64_t *stats);

int32_t ora_cho_lk_casscf_scr(const ora_chol *ch, const double *dlt, double *flt,
                              const double *msq, double *w_pwxy, double factxi,
                              int32_t nfiorb, int32_t naorb, int32_t nchm,
                              const double *ash1, const double *ash2, int32_t do_active,
                              int32_t nscreen, double dmpk, double dfmat, const double *cmo,
                              int32_t nfro, int32_t norb, double exfac, const ora_lk *lk,
                              int64_t nvec_batch, int64_t *stats);

int32_t ora_cd_incore(double *x, int32_t n, double *vec, int32_t mxvec,
                      int32_t *numcho, double thr);

double ora_energy(int64_t np, const double *h, const double *dlt,
                  const double *g, double enuc);

#ifdef __cplusplus
}
#endif
#endif
