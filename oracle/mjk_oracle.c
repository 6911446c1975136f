/*
 * CPU FP64 restatement of OpenMolcas's Cholesky J/K Fock build.
 *
 * TEST INFRASTRUCTURE ONLY: the parity reference for the HIP kernels and the
 * CPU baseline timed beside them.  Nothing in the product path links it.
 *
 * Each function follows the reference routine named in its comment, step for
 * step where the step order fixes the arithmetic (batching, the 1/2 factor on
 * diagonal pairs, the lower-triangle output of DGEMM_Tri).  The BLAS calls of
 * the reference are restated as plain loops (ora_dgemv); OpenMP is used where
 * the reference would use a threaded BLAS, never changing the summation order
 * inside one dot product.  For timing only, ora_set_dgemv routes the DGEMV
 * calls of ora_dgemv and ora_dgemm_tri_tn through a caller-supplied Fortran
 * BLAS dgemv (the reference links one); the checks use the plain loops.
 */
#include "mjk_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* Threads of the OpenMP loops (the threaded BLAS of the reference); the CPU
 * baseline times the restatement on all host cores and on one. */
void ora_set_threads(int32_t n) { omp_set_num_threads(n > 0 ? n : 1); }
int32_t ora_get_threads(void) { return (int32_t)omp_get_max_threads(); }

/* src/misc_util/index_functions.F90:138-146 (1-based, symmetric in i,j). */
int64_t ora_itri(int64_t i, int64_t j) {
  return (j > i) ? j * (j - 1) / 2 + i : i * (i - 1) / 2 + j;
}

/* Inverse of the 0-based LT index: idx = a*(a+1)/2 + b, a >= b. */
static void lt_pair(int64_t idx, int64_t *a, int64_t *b) {
  int64_t r = (int64_t)((sqrt(8.0 * (double)idx + 1.0) - 1.0) * 0.5);
  while (r * (r + 1) / 2 > idx) --r;
  while ((r + 1) * (r + 2) / 2 <= idx) ++r;
  *a = r;
  *b = idx - r * (r + 1) / 2;
}

/* The pair (a, b) of every reduced-set entry, decoded once per call: the
 * reference looks it up in its precomputed iRS2F map
 * (src/cholesky_util/cho_vtra.F90:141-142); pab[2r], pab[2r+1].  Caller frees. */
static int32_t *pair_map(int64_t nrs, const int64_t *rs2lt) {
  int32_t *pab = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)(nrs > 0 ? nrs : 1));
  if (!pab) abort(); /* the checker cannot continue without its map */
  for (int64_t r = 0; r < nrs; ++r) {
    int64_t a, b;
    lt_pair(rs2lt[r], &a, &b);
    pab[2 * r] = (int32_t)a;
    pab[2 * r + 1] = (int32_t)b;
  }
  return pab;
}

/* src/misc_util/fold.F90:14-72 (nSym = 1): square -> LT, off-diagonal x2.
 * A(i,j) column-major; B(iTri(i,j)) for j <= i. */
void ora_fold(int32_t n, const double *a_sq, double *b_lt) {
  int64_t off = 0;
  for (int64_t i = 0; i < n; ++i) {
    for (int64_t j = 0; j < i; ++j) b_lt[off + j] = 2.0 * a_sq[i * n + j];
    b_lt[off + i] = a_sq[i * n + i];
    off += i + 1;
  }
}

/* src/misc_util/unfold.F90:14-80 (nSym = 1): LT -> square, off-diagonal x1/2. */
void ora_unfold(int32_t n, const double *a_lt, double *b_sq) {
  for (int64_t i = 0; i < n; ++i) {
    for (int64_t j = 0; j < i; ++j) {
      double v = 0.5 * a_lt[i * (i + 1) / 2 + j];
      b_sq[j + i * n] = v;
      b_sq[i + j * n] = v;
    }
    b_sq[i + i * n] = a_lt[i * (i + 1) / 2 + i];
  }
}

/* src/fock_util/swap_full2rs.F90:37-52: Xab(jRab,jDen) = XLT(jDen)(iab). */
void ora_swap_full2rs(int64_t nrs, const int64_t *rs2lt, int32_t nmat,
                      int64_t ldlt, const double *xlt, double *xrs) {
  for (int64_t r = 0; r < nrs; ++r)
    for (int32_t d = 0; d < nmat; ++d)
      xrs[r + d * nrs] = xlt[d * ldlt + rs2lt[r]];
}

/* src/fock_util/swap_rs2full.F90:33-58: XLT(jDen)(iab) += Xab(jRab,jDen);
 * with add == 0 the targets are zeroed first. */
void ora_swap_rs2full(int64_t nrs, const int64_t *rs2lt, int32_t nmat,
                      int64_t ldlt, const double *xrs, double *xlt,
                      int32_t add) {
  if (!add)
    for (int32_t d = 0; d < nmat; ++d) memset(xlt + d * ldlt, 0, sizeof(double) * ldlt);
  for (int64_t r = 0; r < nrs; ++r)
    for (int32_t d = 0; d < nmat; ++d) xlt[d * ldlt + rs2lt[r]] += xrs[r + d * nrs];
}

/* Fortran BLAS dgemv (32-bit integers), or NULL for the plain loops. */
typedef void (*ora_dgemv_fn)(const char *trans, const int *m, const int *n, const double *alpha,
                             const double *a, const int *lda, const double *x, const int *incx,
                             const double *beta, double *y, const int *incy);
static ora_dgemv_fn g_dgemv = NULL;

void ora_set_dgemv(void *fn) { g_dgemv = (ora_dgemv_fn)fn; }

static int fits_int(int64_t v) { return v >= 0 && v <= 2147483647; }

/* BLAS DGEMV semantics (column-major A(m,n)), as called through
 * src/misc_util/blas_wrappers.F90:219-293.  beta == 0 overwrites y. */
void ora_dgemv(char trans, int64_t m, int64_t n, double alpha,
               const double *a, int64_t lda, const double *x, double beta,
               double *y) {
  if (g_dgemv && fits_int(m) && fits_int(n) && fits_int(lda) && m > 0 && n > 0) {
    const int im = (int)m, in = (int)n, il = (int)lda, one = 1;
    g_dgemv(&trans, &im, &in, &alpha, a, &il, x, &one, &beta, y, &one);
    return;
  }
  if (trans == 'T' || trans == 't') {
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j) {
      const double *col = a + j * lda;
      double s = 0.0;
      for (int64_t i = 0; i < m; ++i) s += col[i] * x[i];
      y[j] = (beta == 0.0 ? 0.0 : beta * y[j]) + alpha * s;
    }
  } else {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i) {
      double s = 0.0;
      for (int64_t j = 0; j < n; ++j) s += a[i + j * lda] * x[j];
      y[i] = (beta == 0.0 ? 0.0 : beta * y[i]) + alpha * s;
    }
  }
}

/* src/linalg_util/dgemm_tri.F90:46-149, TransA='T', TransB='N':
 * C_LT(i<=j at column j) = beta*C + alpha * sum_l A(l,i) B(l,j),
 * computed column by column as DGEMV('T',k,j,...) (:133-137). */
void ora_dgemm_tri_tn(int64_t n, int64_t k, double alpha, const double *a,
                      int64_t lda, const double *b, int64_t ldb, double beta,
                      double *c_lt) {
  int64_t np = n * (n + 1) / 2;
  if (n == 0) return;
  int noab = (alpha == 0.0) || (k == 0);
  if (noab && beta == 1.0) return;
  if (beta == 0.0)
    memset(c_lt, 0, sizeof(double) * np);
  else if (beta != 1.0)
    for (int64_t i = 0; i < np; ++i) c_lt[i] *= beta;
  if (noab) return;
  if (g_dgemv && fits_int(k) && fits_int(n) && fits_int(lda)) {
    /* dgemm_tri.F90:133-137: column j of C_LT += alpha A(:,1:j)^T B(:,j) */
    const int ik = (int)k, il = (int)lda, one = 1;
    const double b1 = 1.0;
    for (int64_t j = 0; j < n; ++j) {
      const int ij = (int)(j + 1);
      g_dgemv("T", &ik, &ij, &alpha, a, &il, b + j * ldb, &one, &b1, c_lt + j * (j + 1) / 2, &one);
    }
    return;
  }
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t j = 0; j < n; ++j) {
    double *cj = c_lt + j * (j + 1) / 2;
    const double *bj = b + j * ldb;
    for (int64_t i = 0; i <= j; ++i) {
      const double *ai = a + i * lda;
      double s = 0.0;
      for (int64_t l = 0; l < k; ++l) s += ai[l] * bj[l];
      cj[i] += alpha * s;
    }
  }
}

/* src/linalg_util/dgemm_tri.F90:126-130, TransA='N', TransB='T':
 * C_LT(i<=j) = beta*C + alpha * sum_l A(i,l) B(j,l)  (A(n,k), B(n,k)). */
void ora_dgemm_tri_nt(int64_t n, int64_t k, double alpha, const double *a,
                      int64_t lda, const double *b, int64_t ldb, double beta,
                      double *c_lt) {
  int64_t np = n * (n + 1) / 2;
  if (n == 0) return;
  int noab = (alpha == 0.0) || (k == 0);
  if (noab && beta == 1.0) return;
  if (beta == 0.0)
    memset(c_lt, 0, sizeof(double) * np);
  else if (beta != 1.0)
    for (int64_t i = 0; i < np; ++i) c_lt[i] *= beta;
  if (noab) return;
  for (int64_t j = 0; j < n; ++j) {
    double *cj = c_lt + j * (j + 1) / 2;
    for (int64_t i = 0; i <= j; ++i) {
      double s = 0.0;
      for (int64_t l = 0; l < k; ++l) s += a[i + l * lda] * b[j + l * ldb];
      cj[i] += alpha * s;
    }
  }
}

/* src/cholesky_util/cho_vtra.F90:318-363 (iSwap = 2, JSYM = 1) after the
 * zeroing in src/cholesky_util/cho_x_getvtra.F90:77-79:
 *   ChoT(p,J,b) = sum_a xfd * L_ab,J * C(p,a), both (a,b) and (b,a),
 * xfd = 1/2 on the diagonal.  chot is A3(norb, jnum, n); mo is A2(norb, n). */
void ora_cho_vtra2(int32_t n, int64_t nrs, const int64_t *rs2lt,
                   int64_t jnum, const double *lrs, int32_t norb,
                   const double *mo, double *chot) {
  memset(chot, 0, sizeof(double) * (size_t)norb * jnum * n);
  int32_t *pab = pair_map(nrs, rs2lt);
#pragma omp parallel for schedule(static)
  for (int64_t jv = 0; jv < jnum; ++jv) {
    const double *l = lrs + jv * nrs;
    for (int64_t r = 0; r < nrs; ++r) {
      const int64_t a = pab[2 * r], b = pab[2 * r + 1];
      double xfd = (a == b) ? 0.5 : 1.0;
      double s = xfd * l[r];
      double *xa = chot + (size_t)norb * (jv + jnum * a);
      double *xb = chot + (size_t)norb * (jv + jnum * b);
      const double *ca = mo + (size_t)norb * a;
      const double *cb = mo + (size_t)norb * b;
      for (int32_t p = 0; p < norb; ++p) xa[p] += s * cb[p];
      for (int32_t p = 0; p < norb; ++p) xb[p] += s * ca[p];
    }
  }
  free(pab);
}

/* src/cholesky_util/cho_vtra.F90:130-163 (iSwap = 0, JSYM = 1):
 * ChoT(p,b,J) = sum_a xfd * L_ab,J * C(p,a); chot is A3(norb, n, jnum). */
void ora_cho_vtra0(int32_t n, int64_t nrs, const int64_t *rs2lt,
                   int64_t jnum, const double *lrs, int32_t norb,
                   const double *mo, double *chot) {
  memset(chot, 0, sizeof(double) * (size_t)norb * jnum * n);
  int32_t *pab = pair_map(nrs, rs2lt);
#pragma omp parallel for schedule(static)
  for (int64_t jv = 0; jv < jnum; ++jv) {
    const double *l = lrs + jv * nrs;
    double *xj = chot + (size_t)norb * n * jv;
    for (int64_t r = 0; r < nrs; ++r) {
      const int64_t a = pab[2 * r], b = pab[2 * r + 1];
      double xfd = (a == b) ? 0.5 : 1.0;
      double s = xfd * l[r];
      double *xa = xj + (size_t)norb * a;
      double *xb = xj + (size_t)norb * b;
      const double *ca = mo + (size_t)norb * a;
      const double *cb = mo + (size_t)norb * b;
      for (int32_t p = 0; p < norb; ++p) xa[p] += s * cb[p];
      for (int32_t p = 0; p < norb; ++p) xb[p] += s * ca[p];
    }
  }
  free(pab);
}

/* Copy vectors jv0..jv0+jnum-1 (all of one reduced set) into Lrs(nrs,jnum):
 * the role of CHO_VECRD (src/cholesky_util/cho_vecrd.F90:12-64) reading from
 * the in-memory buffer.  Returns the number of vectors delivered. */
static int64_t vec_read(const ora_chol *ch, int64_t jv0, int64_t jnum,
                        int64_t nrs, double *lrs) {
  for (int64_t j = 0; j < jnum; ++j) {
    memcpy(lrs + j * nrs, ch->lvec + ch->voff[jv0 + j], sizeof(double) * nrs);
  }
  return jnum;
}

/* First vector and count of reduced set r: Cho_X_nVecRS
 * (src/cholesky_util/cho_x_nvecrs.F90:37-126). */
static void nvec_rs(const ora_chol *ch, int32_t r, int64_t *iv, int64_t *nv) {
  *iv = -1;
  *nv = 0;
  for (int64_t j = 0; j < ch->nvec; ++j) {
    if (ch->vred[j] == r) {
      if (*iv < 0) *iv = j;
      ++*nv;
    }
  }
}

/* src/fock_util/cho_fscf.F90:14-371 with nSym = 1.
 *   FLT(d) += J[DLT(1)] - ExFac * sum_{J,k} X(k,J,a) X(k,J,b)
 * porb[d] is Porb(d)%SB(1)%A2(nforb[d]+niorb[d], n).  nvec_batch plays the
 * role of nVec (the vectors that fit in memory, :145).  Returns rc. */
int32_t ora_cho_fscf(const ora_chol *ch, int32_t nden, double *flt,
                     const int32_t *nforb, const int32_t *niorb,
                     const double *const *porb, const double *dlt,
                     double exfac, int64_t nvec_batch) {
  if (nden != 1 && nden != 2) return 99; /* :75-78 abend */
  const int32_t n = ch->n;
  const int64_t np = (int64_t)n * (n + 1) / 2;
  const double fact_ci = 1.0, fact_xi = -exfac; /* :64-65 */
  int32_t maxocc = 0;
  for (int d = 0; d < nden; ++d)
    if (nforb[d] + niorb[d] > maxocc) maxocc = nforb[d] + niorb[d];

  for (int32_t jred = 0; jred < ch->nred; ++jred) {
    int64_t ivrs, nvrs;
    nvec_rs(ch, jred, &ivrs, &nvrs);
    if (nvrs == 0) continue;
    const int64_t nrs = ch->nrs[jred];
    const int64_t *map = ch->rs2lt + ch->rs_off[jred];
    int64_t nvec = nvec_batch < nvrs ? nvec_batch : nvrs;
    if (nvec < 1) return 33; /* :147-155 insufficient memory */
    double *drs = calloc((size_t)nrs, sizeof(double));
    double *frs = calloc((size_t)nrs, sizeof(double));
    double *lrs = malloc(sizeof(double) * (size_t)nrs * nvec);
    double *vj = malloc(sizeof(double) * (size_t)nvec);
    double *laq = malloc(sizeof(double) * (size_t)(maxocc > 0 ? maxocc : 1) * nvec * n);
    ora_swap_full2rs(nrs, map, 1, np, dlt, drs); /* :163-168 */
    int64_t nbatch = (nvrs - 1) / nvec + 1;
    for (int64_t ib = 0; ib < nbatch; ++ib) {
      int64_t jnum = (ib == nbatch - 1) ? nvrs - nvec * (nbatch - 1) : nvec;
      int64_t jvec = nvec * ib + ivrs;
      int64_t numv = vec_read(ch, jvec, jnum, nrs, lrs);
      if (numv <= 0 || numv != jnum) { /* :193-196 */
        free(drs); free(frs); free(lrs); free(vj); free(laq);
        return 77;
      }
      /* Coulomb, :200-222 */
      ora_dgemv('T', nrs, jnum, 1.0, lrs, nrs, drs, 0.0, vj);
      double fact = (jvec - ivrs) < 1 ? (double)(jvec - ivrs) : 1.0;
      ora_dgemv('N', nrs, jnum, fact_ci, lrs, nrs, vj, fact, frs);
      /* Exchange, :226-289 */
      for (int32_t d = 0; d < nden; ++d) {
        int32_t nk = nforb[d] + niorb[d];
        if (nk == 0) continue; /* iSkip, :238-245 */
        ora_cho_vtra2(n, nrs, map, jnum, lrs, nk, porb[d], laq);
        ora_dgemm_tri_tn(n, (int64_t)nk * jnum, fact_xi, laq, (int64_t)nk * jnum,
                         laq, (int64_t)nk * jnum, 1.0, flt + d * np);
      }
    }
    for (int32_t d = 0; d < nden; ++d) /* :299-305 */
      ora_swap_rs2full(nrs, map, 1, np, frs, flt + d * np, 1);
    free(drs); free(frs); free(lrs); free(vj); free(laq);
  }
  return 0;
}

/* src/fock_util/cho_fock_dft_red.F90:14-235 (JSYM = 1): Coulomb only.
 *   Frs = sum_J L_rs,J V_J per reduced set, V_J = sum_rs L_rs,J Drs
 * written into FLT for the first reduced set (add = JRED > JRED1, :190-191)
 * and added for the later ones.  Returns irc. */
int32_t ora_cho_fock_dft_red(const ora_chol *ch, const double *dlt, double *flt,
                             int64_t nvec_batch) {
  const int32_t n = ch->n;
  const int64_t np = (int64_t)n * (n + 1) / 2;
  if (ch->nvec < 1) return 0; /* :60 */
  const int32_t jred1 = ch->vred[0], jred2 = ch->vred[ch->nvec - 1]; /* :70-71 */
  for (int32_t jred = jred1; jred <= jred2; ++jred) {
    int64_t ivrs, nvrs;
    nvec_rs(ch, jred, &ivrs, &nvrs);
    if (nvrs == 0) continue;
    const int64_t nrs = ch->nrs[jred];
    const int64_t *map = ch->rs2lt + ch->rs_off[jred];
    int64_t nvec = nvec_batch < nvrs ? nvec_batch : nvrs;
    if (nvec < 1) return 33; /* :104-111 */
    double *drs = calloc((size_t)nrs, sizeof(double));
    double *frs = calloc((size_t)nrs, sizeof(double));
    double *lrs = malloc(sizeof(double) * (size_t)nrs * nvec);
    double *vj = malloc(sizeof(double) * (size_t)nvec);
    ora_swap_full2rs(nrs, map, 1, np, dlt, drs); /* :122 */
    int64_t nbatch = (nvrs - 1) / nvec + 1;
    for (int64_t ib = 0; ib < nbatch; ++ib) {
      int64_t jnum = (ib == nbatch - 1) ? nvrs - nvec * (nbatch - 1) : nvec;
      int64_t jvec = nvec * ib + ivrs;
      int64_t numv = vec_read(ch, jvec, jnum, nrs, lrs);
      if (numv <= 0 || numv != jnum) {
        free(drs); free(frs); free(lrs); free(vj);
        return 77;
      }
      ora_dgemv('T', nrs, jnum, 1.0, lrs, nrs, drs, 0.0, vj); /* :167 */
      double xfac = (jvec - ivrs) < 1 ? (double)(jvec - ivrs) : 1.0;
      ora_dgemv('N', nrs, jnum, 1.0, lrs, nrs, vj, xfac, frs); /* :174 */
    }
    ora_swap_rs2full(nrs, map, 1, np, frs, flt, jred > jred1); /* :188-192 */
    free(drs); free(frs); free(lrs); free(vj);
  }
  return 0;
}

/* src/fock_util/cho_lk_scf.F90:14-1103 with nSym = 1 and every shell pair and
 * orbital kept (no LK screening: the limit dmpk -> 0 of the reference).
 *   FLT(d) += J[PLT(1)] + KLT(d),  KLT(d) += -FactXI * sum_{J,k} L(a,J)[k] L(b,J)[k]
 * with L(a,J)[k] = sum_b L^J_ab C(b,k) over the full square L^J
 * (CHO_GetShFull + DGEMV, :507,:675-689) and KLT(d) in the reference's
 * shell-pair-blocked storage (:274-292): pairs iaSh >= ibSh in order, an
 * off-diagonal block nBasSh(iaSh) x nBasSh(ibSh) column-major, a diagonal
 * block lower-packed; folded into FLT at the end (:963-1028).
 * porb[d] is Porb(d)%SB(1)%A2(n,n), AO index fastest, columns 0..nforb+niorb-1
 * used.  klt (nden blocks of n(n+1)/2, may be NULL) is accumulated.  nbassh
 * lists the shell sizes (sum = n).  Returns rc. */
int32_t ora_cho_lk_scf(const ora_chol *ch, int32_t nden, double *flt, double *klt,
                       const int32_t *nforb, const int32_t *niorb,
                       const double *const *porb, const double *plt, double factxi,
                       int32_t nshell, const int32_t *nbassh, int64_t nvec_batch) {
  if (nden != 1 && nden != 2) return 99; /* :93-96 abend */
  const int32_t n = ch->n;
  const int64_t np = (int64_t)n * (n + 1) / 2;
  int64_t tot = 0;
  for (int32_t s = 0; s < nshell; ++s) tot += nbassh[s];
  if (nshell < 1 || tot != n) return 99;
  /* shell offsets kOffSh and shell-pair offsets nnBfShp (:255-292) */
  int64_t *koff = malloc(sizeof(int64_t) * (size_t)nshell);
  int64_t *poff = malloc(sizeof(int64_t) * (size_t)nshell * (nshell + 1) / 2);
  int64_t lk = 0, ao = 0;
  for (int32_t a = 0; a < nshell; ++a) {
    koff[a] = ao;
    ao += nbassh[a];
    for (int32_t b = 0; b <= a; ++b) {
      poff[(int64_t)a * (a + 1) / 2 + b] = lk;
      lk += (int64_t)nbassh[a] * nbassh[b] -
            (a == b ? (int64_t)nbassh[a] * (nbassh[a] - 1) / 2 : 0);
    }
  }
  double *kbuf = klt ? klt : calloc((size_t)(nden * np), sizeof(double));
  for (int32_t jred = 0; jred < ch->nred; ++jred) {
    int64_t ivrs, nvrs;
    nvec_rs(ch, jred, &ivrs, &nvrs);
    if (nvrs == 0) continue;
    const int64_t nrs = ch->nrs[jred];
    const int64_t *map = ch->rs2lt + ch->rs_off[jred];
    int64_t nvec = nvec_batch < nvrs ? nvec_batch : nvrs;
    if (nvec < 1) { free(koff); free(poff); if (!klt) free(kbuf); return 33; }
    double *drs = calloc((size_t)nrs, sizeof(double));
    double *frs = calloc((size_t)nrs, sizeof(double));
    double *lrs = malloc(sizeof(double) * (size_t)nrs * nvec);
    double *vj = malloc(sizeof(double) * (size_t)nvec);
    double *lfull = malloc(sizeof(double) * (size_t)nvec * n * n);
    double *lab = malloc(sizeof(double) * (size_t)nvec * n);
    ora_swap_full2rs(nrs, map, 1, np, plt, drs); /* :397-403 */
    int64_t nbatch = (nvrs - 1) / nvec + 1;
    for (int64_t ib = 0; ib < nbatch; ++ib) {
      int64_t jnum = (ib == nbatch - 1) ? nvrs - nvec * (nbatch - 1) : nvec;
      int64_t jvec = nvec * ib + ivrs;
      int64_t numv = vec_read(ch, jvec, jnum, nrs, lrs);
      if (numv <= 0 || numv != jnum) { /* :419-424 */
        free(drs); free(frs); free(lrs); free(vj); free(lfull); free(lab);
        free(koff); free(poff); if (!klt) free(kbuf);
        return 77;
      }
      /* Coulomb of the total density, :428-449 */
      ora_dgemv('T', nrs, jnum, 1.0, lrs, nrs, drs, 0.0, vj);
      double fact = (jvec - ivrs) < 1 ? (double)(jvec - ivrs) : 1.0;
      ora_dgemv('N', nrs, jnum, 1.0, lrs, nrs, vj, fact, frs);
      /* full square L(a,b) of each vector (CHO_GetShFull, :507) */
      memset(lfull, 0, sizeof(double) * (size_t)jnum * n * n);
      for (int64_t j = 0; j < jnum; ++j) {
        double *lf = lfull + j * n * n;
        for (int64_t r = 0; r < nrs; ++r) {
          int64_t a, b;
          lt_pair(map[r], &a, &b);
          lf[a * n + b] = lrs[r + j * nrs];
          lf[b * n + a] = lrs[r + j * nrs];
        }
      }
      for (int32_t d = 0; d < nden; ++d) {
        const int32_t nk = nforb[d] + niorb[d];
        double *kd = kbuf + d * np;
        for (int32_t k = 0; k < nk; ++k) {
          const double *ck = porb[d] + (int64_t)n * k;
          /* L(a,J)[k] = sum_b L(a,b) C(b,k), :675-689 */
          for (int64_t j = 0; j < jnum; ++j)
            for (int64_t a = 0; a < n; ++a) {
              const double *row = lfull + j * n * n + a * n;
              double s = 0.0;
              for (int64_t b = 0; b < n; ++b) s += row[b] * ck[b];
              lab[j * n + a] = s;
            }
          /* shell-pair blocks, DGEMM_Tri on diagonal pairs (:803), DGEMM off it (:830) */
          for (int32_t sa = 0; sa < nshell; ++sa)
            for (int32_t sb = 0; sb <= sa; ++sb) {
              const int64_t off = poff[(int64_t)sa * (sa + 1) / 2 + sb];
              for (int64_t bl = 0; bl < nbassh[sb]; ++bl)
                for (int64_t al = (sa == sb ? bl : 0); al < nbassh[sa]; ++al) {
                  const int64_t ag = koff[sa] + al, bg = koff[sb] + bl;
                  double s = 0.0;
                  for (int64_t j = 0; j < jnum; ++j) s += lab[j * n + ag] * lab[j * n + bg];
                  const int64_t idx = sa == sb ? al * (al + 1) / 2 + bl : nbassh[sa] * bl + al;
                  kd[off + idx] += -factxi * s;
                }
            }
        }
      }
    }
    for (int32_t d = 0; d < nden; ++d) /* :912-918 */
      ora_swap_rs2full(nrs, map, 1, np, frs, flt + d * np, 1);
    free(drs); free(frs); free(lrs); free(vj); free(lfull); free(lab);
  }
  /* KLT -> FLT (:963-1028) */
  for (int32_t d = 0; d < nden; ++d) {
    const double *kd = kbuf + d * np;
    double *fd = flt + d * np;
    for (int32_t sa = 0; sa < nshell; ++sa)
      for (int32_t sb = 0; sb <= sa; ++sb) {
        const int64_t off = poff[(int64_t)sa * (sa + 1) / 2 + sb];
        for (int64_t bl = 0; bl < nbassh[sb]; ++bl)
          for (int64_t al = (sa == sb ? bl : 0); al < nbassh[sa]; ++al) {
            const int64_t ag = koff[sa] + al, bg = koff[sb] + bl;
            const int64_t idx = sa == sb ? al * (al + 1) / 2 + bl : nbassh[sa] * bl + al;
            fd[ag * (ag + 1) / 2 + bg] += kd[off + idx];
          }
      }
  }
  free(koff); free(poff);
  if (!klt) free(kbuf);
  return 0;
}

/* src/fock_util/cho_fmcscf.F90:14-585 with nSym = 1, followed by
 * src/fock_util/cho_eval_waxy.F90:12-150 for the (pw|xy) integrals.
 *   FI += J[DI] + FactXI * sum X_I X_I^T     (POrb(1), nforb+niorb rows)
 *   FA += J[DA] - 1/2   * sum X_A X_A^T      (POrb(2), nchm rows)
 *   W_PWXY(p,w,xy) = sum_J (pw|J)(xy|J), p over norb MOs after nfro frozen.
 * dlt is DLT(1:2), flt is FLT(1:2), each np long.  cmo is CMO(n,n)
 * column-major.  w_pwxy receives norb*naorb*naorb(naorb+1)/2 values. */
static int32_t fmcscf_body(const ora_chol *ch, double *flt, int32_t nforb,
                           int32_t niorb, int32_t naorb, double factxi,
                           double fact_xa, const double *dlt, int32_t do_active,
                           const double *porb1, int32_t nchm,
                           const double *porb2, const double *porb3,
                           double *w_pwxy, const double *cmo, int32_t nfro,
                           int32_t norb, int64_t nvec_batch);

int32_t ora_cho_fmcscf(const ora_chol *ch, double *flt, int32_t nforb,
                       int32_t niorb, int32_t naorb, double factxi,
                       const double *dlt, int32_t do_active,
                       const double *porb1, int32_t nchm,
                       const double *porb2, const double *porb3,
                       double *w_pwxy, const double *cmo, int32_t nfro,
                       int32_t norb, double exfac, int64_t nvec_batch) {
  (void)exfac; /* only triggers a warning message in the reference, :88-94 */
  return fmcscf_body(ch, flt, nforb, niorb, naorb, factxi, -0.5 /* FactXA, :74 */, dlt,
                     do_active, porb1, nchm, porb2, porb3, w_pwxy, cmo, nfro, norb,
                     nvec_batch);
}

/* src/fock_util/cho_lk_casscf.F90:14-1268 with nSym = 1 and no screening:
 * the same FI, FA and (tw|xy) as CHO_FMCSCF with the exchange factors
 * FactX = [FactXI*ExFac, -1/2*ExFac] (:113).  msq (the inactive orbitals or
 * DI vectors) and ash2 (the DA vectors) are square A2(n,n), AO index fastest,
 * first nfiorb / nchm columns used; ash1 holds the active MOs orbital-major
 * A2(naorb,n). */
int32_t ora_cho_lk_casscf(const ora_chol *ch, const double *dlt, double *flt,
                          const double *msq, double *w_pwxy, double factxi,
                          int32_t nfiorb, int32_t naorb, int32_t nchm,
                          const double *ash1, const double *ash2, int32_t do_active,
                          const double *cmo, int32_t nfro, int32_t norb, double exfac,
                          int64_t nvec_batch) {
  const int32_t n = ch->n;
  double *p1 = calloc((size_t)(nfiorb > 0 ? nfiorb : 1) * n, sizeof(double));
  double *p2 = calloc((size_t)(nchm > 0 ? nchm : 1) * n, sizeof(double));
  for (int64_t k = 0; k < nfiorb; ++k)
    for (int64_t b = 0; b < n; ++b) p1[k + nfiorb * b] = msq[b + (int64_t)n * k];
  for (int64_t k = 0; k < nchm; ++k)
    for (int64_t b = 0; b < n; ++b) p2[k + nchm * b] = ash2[b + (int64_t)n * k];
  int32_t rc = fmcscf_body(ch, flt, 0, nfiorb, naorb, factxi * exfac, -0.5 * exfac, dlt,
                           do_active, p1, nchm, p2, ash1, w_pwxy, cmo, nfro, norb,
                           nvec_batch);
  free(p1);
  free(p2);
  return rc;
}

static int32_t fmcscf_body(const ora_chol *ch, double *flt, int32_t nforb,
                           int32_t niorb, int32_t naorb, double factxi,
                           double fact_xa, const double *dlt, int32_t do_active,
                           const double *porb1, int32_t nchm,
                           const double *porb2, const double *porb3,
                           double *w_pwxy, const double *cmo, int32_t nfro,
                           int32_t norb, int64_t nvec_batch) {
  const int32_t n = ch->n;
  const int64_t np = (int64_t)n * (n + 1) / 2;
  const double fact_ci = 1.0, fact_ca = 1.0; /* :71 */
  const int32_t nden = do_active ? 2 : 1;
  const int32_t nk = nforb + niorb;
  const int64_t nxy = (int64_t)naorb * (naorb + 1) / 2;
  /* Scr(w,a,xy), the twxy buffer accumulated over all batches. */
  double *scr = calloc((size_t)naorb * n * (nxy > 0 ? nxy : 1), sizeof(double));
  int64_t total_vec = 0;
  for (int32_t r = 0; r < ch->nred; ++r) {
    int64_t iv, nv;
    nvec_rs(ch, r, &iv, &nv);
    total_vec += nv;
  }
  (void)total_vec;
  int32_t last_red = -1;
  for (int32_t r = 0; r < ch->nred; ++r) {
    int64_t iv, nv;
    nvec_rs(ch, r, &iv, &nv);
    if (nv > 0) last_red = r;
  }

  for (int32_t jred = 0; jred < ch->nred; ++jred) {
    int64_t ivrs, nvrs;
    nvec_rs(ch, jred, &ivrs, &nvrs);
    if (nvrs == 0) continue;
    const int64_t nrs = ch->nrs[jred];
    const int64_t *map = ch->rs2lt + ch->rs_off[jred];
    int64_t nvec = nvec_batch < nvrs ? nvec_batch : nvrs;
    if (nvec < 1) { free(scr); return 33; }
    double *drs = calloc((size_t)nrs * 2, sizeof(double));
    double *frs = calloc((size_t)nrs * 2, sizeof(double));
    double *lrs = malloc(sizeof(double) * (size_t)nrs * nvec);
    double *vj = malloc(sizeof(double) * (size_t)nvec);
    int32_t mx = nk > nchm ? nk : nchm;
    if (naorb > mx) mx = naorb;
    double *laq = malloc(sizeof(double) * (size_t)(mx > 0 ? mx : 1) * nvec * n);
    double *lxy = malloc(sizeof(double) * (size_t)(nxy > 0 ? nxy : 1) * nvec);
    ora_swap_full2rs(nrs, map, nden, np, dlt, drs); /* :214-218 */
    int64_t nbatch = (nvrs - 1) / nvec + 1;
    for (int64_t ib = 0; ib < nbatch; ++ib) {
      int64_t jnum = (ib == nbatch - 1) ? nvrs - nvec * (nbatch - 1) : nvec;
      int64_t jvec = nvec * ib + ivrs;
      int64_t numv = vec_read(ch, jvec, jnum, nrs, lrs);
      if (numv <= 0 || numv != jnum) {
        free(drs); free(frs); free(lrs); free(vj); free(laq); free(lxy); free(scr);
        return 77;
      }
      double fact = (jvec - ivrs) < 1 ? (double)(jvec - ivrs) : 1.0;
      /* inactive and active Coulomb, :250-291 */
      ora_dgemv('T', nrs, jnum, 1.0, lrs, nrs, drs, 0.0, vj);
      ora_dgemv('N', nrs, jnum, fact_ci, lrs, nrs, vj, fact, frs);
      if (do_active) {
        ora_dgemv('T', nrs, jnum, 1.0, lrs, nrs, drs + nrs, 0.0, vj);
        ora_dgemv('N', nrs, jnum, fact_ca, lrs, nrs, vj, fact, frs + nrs);
      }
      /* inactive exchange, :299-341 */
      if (nk > 0) {
        ora_cho_vtra2(n, nrs, map, jnum, lrs, nk, porb1, laq);
        ora_dgemm_tri_tn(n, (int64_t)nk * jnum, factxi, laq, (int64_t)nk * jnum,
                         laq, (int64_t)nk * jnum, 1.0, flt);
      }
      /* active exchange with the Cholesky MOs of DA, :343-401 */
      if (do_active && nchm > 0) {
        ora_cho_vtra2(n, nrs, map, jnum, lrs, nchm, porb2, laq);
        ora_dgemm_tri_tn(n, (int64_t)nchm * jnum, fact_xa, laq, (int64_t)nchm * jnum,
                         laq, (int64_t)nchm * jnum, 1.0, flt + np);
      }
      /* (wa|xy): L(v,b,J) then Lvw,J in LT, :403-449 */
      if (naorb > 0) {
        ora_cho_vtra0(n, nrs, map, jnum, lrs, naorb, porb3, laq);
        for (int64_t jc = 0; jc < jnum; ++jc)
          ora_dgemm_tri_nt(naorb, n, 1.0, laq + (size_t)naorb * n * jc, naorb,
                           porb3, naorb, 0.0, lxy + nxy * jc);
        /* CHO_eval_waxy accumulation, src/fock_util/cho_eval_waxy.F90:51-72:
         * Scr(wa,xy) += sum_J L(wa,J) Lxy(xy,J). */
        const int64_t nwa = (int64_t)naorb * n;
        for (int64_t xy = 0; xy < nxy; ++xy) {
          double *sc = scr + nwa * xy;
          for (int64_t jc = 0; jc < jnum; ++jc) {
            double f = lxy[nxy * jc + xy];
            const double *lw = laq + nwa * jc;
            for (int64_t i = 0; i < nwa; ++i) sc[i] += lw[i] * f;
          }
        }
      }
    }
    ora_swap_rs2full(nrs, map, nden, np, frs, flt, 1); /* :508-512 */
    free(drs); free(frs); free(lrs); free(vj); free(laq); free(lxy);
  }

  /* DoTraInt at the last batch of the last reduced set
   * (src/fock_util/cho_eval_waxy.F90:76-147, nSym = 1):
   * W_PWXY(p,w,xy) = sum_a CMO(a,nfro+p) Scr(w,a,xy). */
  if (naorb > 0 && last_red >= 0 && w_pwxy) {
    const int64_t npw = (int64_t)norb * naorb;
    for (int64_t xy = 0; xy < nxy; ++xy) {
      const double *sc = scr + (int64_t)naorb * n * xy;
      for (int32_t w = 0; w < naorb; ++w) {
        for (int32_t p = 0; p < norb; ++p) {
          const double *cp = cmo + (int64_t)n * (nfro + p);
          double s = 0.0;
          for (int32_t a = 0; a < n; ++a) s += cp[a] * sc[w + (int64_t)naorb * a];
          w_pwxy[npw * xy + p + (int64_t)norb * w] = s;
        }
      }
    }
  }
  free(scr);
  return 0;
}

/* src/cholesky_util/cd_incore.F90:35-60 + cd_incore_1.F90:14-105:
 * incomplete pivoted Cholesky of the n-by-n matrix x (column-major, destroyed)
 * into vec(n,mxvec).  rc 101: negative diagonal below ThrFail; 102: too many
 * vectors; -1: mxvec < 1. */
int32_t ora_cd_incore(double *x, int32_t n, double *vec, int32_t mxvec,
                      int32_t *numcho, double thr) {
  const double def_thr = 1.0e-6, thr_fail = -1.0e-8, thr_neg = -1.0e-13;
  *numcho = 0;
  if (n < 1) return 0;
  if (mxvec <= 0) return -1;
  if (thr < 0.0) thr = def_thr;
  const double acc = fmin(1.0e-12, thr * 1.0e-2);
#define X(i, j) x[(i) + (int64_t)(j) * n]
#define V(i, j) vec[(i) + (int64_t)(j) * n]
  for (int32_t pass = 0; pass < n; ++pass) {
    if (X(0, 0) < thr_neg) {
      if (X(0, 0) < thr_fail) return 101;
      for (int32_t i = 0; i < n; ++i) { X(i, 0) = 0.0; X(0, i) = 0.0; }
    }
    double xmax = 0.0;
    int32_t imax = -1;
    for (int32_t i = 0; i < n; ++i) {
      if (X(i, i) < thr_neg) {
        if (X(i, i) < thr_fail) return 101;
        for (int32_t k = 0; k < n; ++k) { X(k, i) = 0.0; X(i, k) = 0.0; }
      }
      if (X(i, i) > xmax + acc) { xmax = X(i, i); imax = i; }
    }
    if (xmax <= thr) return 0;
    const double xfac = 1.0 / sqrt(fabs(xmax));
    if (*numcho == mxvec) return 102;
    for (int32_t j = 0; j < *numcho; ++j) {
      const double vij = V(imax, j);
      for (int32_t i = 0; i < n; ++i) X(i, imax) -= vij * V(i, j);
    }
    X(imax, imax) = xmax;
    const int32_t nc = (*numcho)++;
    for (int32_t i = 0; i < n; ++i) V(i, nc) = (X(i, i) == 0.0) ? 0.0 : xfac * X(i, imax);
    for (int32_t i = 0; i < n; ++i) X(i, i) -= V(i, nc) * V(i, nc);
    X(imax, imax) = 0.0;
  }
#undef X
#undef V
  return 0;
}

/* src/scf/eneclc.F90:58,77,104 (nD = 1): E = h.D + 1/2 G.D + E_nuc, dot
 * products over LT storage with the folded density. */
double ora_energy(int64_t np, const double *h, const double *dlt,
                  const double *g, double enuc) {
  double e1 = 0.0, e2 = 0.0;
  for (int64_t i = 0; i < np; ++i) e1 += h[i] * dlt[i];
  for (int64_t i = 0; i < np; ++i) e2 += g[i] * dlt[i];
  return e1 + 0.5 * e2 + enuc;
}
