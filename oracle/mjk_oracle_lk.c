/*
 * CPU FP64 restatement of the LK ("local K") screened exchange of OpenMolcas's
 * CHO_LK_SCF and CHO_LK_CASSCF (nSym = 1, JSYM = 1).
 *
 * TEST INFRASTRUCTURE ONLY: the parity reference for the screened HIP path.
 * Nothing in the product path links it.
 *
 * The screening decisions follow the reference step for step, because they
 * decide which contributions exist at all:
 *   - thresholds tau, thrv (src/fock_util/cho_lk_scf.F90:143-171,
 *     src/fock_util/cho_lk_casscf.F90:164-200) from the basic threshold of
 *     Cho_LK_ScreeningThreshold (src/fock_util/cho_lk_screeningthreshold.F90:
 *     34-39);
 *   - the diagonal D(ab) read at the start of the call and depleted by the
 *     vectors of every batch (Update, cho_lk_scf.F90:180,834-880), or
 *     re-estimated from the batch (Estimate, :461-479);
 *   - per shell pair the largest squared norm over the batch's vectors,
 *     SvShp (CHO_GetShFull, src/fock_util/cho_getshfull.F90:48-96);
 *   - every nScreen reduced sets (and at the first batch) the per-orbital
 *     shell lists: Y(l)[k] = sum_n sqrt|D(l,n)| |C(n,k)|, its largest element
 *     per shell MLk, sorted by selection with a cut (SCF:
 *     MLk(j)*MLk(1) >= tau, cho_lk_scf.F90:547-627; CASSCF: MLk(j) >=
 *     sqrt(tau), cho_lk_casscf.F90:604-698);
 *   - the half-transform of a listed shell a over the shells b with
 *     sqrt(|SumAClk(b,k) SvShp(ab)|) >= thrv (:659-702), Keep(a) = any b;
 *   - Faa(a) = max_{i in a} sum_J L(i,J)[k]^2 (:713-743) and the K block of a
 *     listed pair a >= b when MLk*MLk >= tau, sqrt|Faa(a)Faa(b)| >= tau/MaxRed
 *     and the Keep flags hold (:749-830), in the shell-pair-blocked KLT
 *     storage (:274-292).
 * Coulomb, (pw|xy) and the KLT -> FLT fold are the unscreened routines of
 * mjk_oracle.c.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mjk_oracle.h"

/* Cho_LK_ScreeningThreshold(delta), cho_lk_screeningthreshold.F90:34-39. */
static double lk_basic_threshold(double thrcom, double delta) {
  double thr = thrcom < 1.0e-6 ? thrcom : 1.0e-6;
  if (delta >= 0.0 && delta <= 1.0) thr *= delta;
  return thr > 1.0e-15 ? thr : 1.0e-15;
}

/* tau(d) = LKThr/max(1,nOrb_d)*dmpk (/MaxRed with Estimate) and
 * thrv(d) = sqrt(LKThr/(max(1,nOrb_d)*NumVT))*dmpk:
 * cho_lk_scf.F90:143-171, cho_lk_casscf.F90:164-200 (same formulas with the
 * inactive / active orbital counts). */
void ora_lk_thresholds(double thrcom, double dfmat, double dmpk, int32_t nden,
                       const int32_t *nocc, int64_t numvt, int32_t estimate,
                       int32_t maxred, double *tau, double *thrv) {
  const double lkthr = lk_basic_threshold(thrcom, dfmat);
  for (int32_t d = 0; d < nden; ++d) {
    const double no = nocc[d] > 1 ? (double)nocc[d] : 1.0;
    tau[d] = (lkthr / no) * dmpk;
    if (estimate) tau[d] = tau[d] / (double)maxred;
    thrv[d] = sqrt(lkthr / (no * (double)numvt)) * dmpk;
  }
}

typedef struct {
  int32_t n, nshell;
  const int32_t *nbassh;
  int64_t *koff;     /* [nshell] first AO of each shell */
  int32_t *shell;    /* [n] shell of each AO */
  int64_t *poff;     /* [nshell(nshell+1)/2] KLT block offsets (nnBfShp) */
  int32_t nden;
  int32_t nocc[2];
  const double *csq[2]; /* A2(n,n), AO index fastest */
  double fact[2];       /* factor of the K block update */
  double tau[2], thrv[2], xtau[2];
  int32_t cas_cut;
  double maxredt;
  int32_t estimate, update, nscreen;
  /* state carried over batches */
  double *diag;   /* [np] */
  double *diah;   /* [n*n] sqrt|D| */
  double *sumac;  /* [nden][maxocc][nshell] */
  double *mlk;    /* [nden][maxocc][nshell] */
  int32_t *indx;  /* [nden][maxocc][nshell + 1]; [0] = numSh */
  int32_t maxocc;
  /* batch scratch */
  double *svshp, *sv2; /* [nshp] */
  int64_t *nrs_shp;    /* [nshp] pairs of the current reduced set per shell pair */
  double *lfull;       /* [jnum][n][n] */
  double *lab;         /* [jnum][n] */
  double *ylk, *absc, *faa;
  unsigned char *keep;
  int64_t *stats;
} lk_state;

static int64_t shp_index(int32_t a, int32_t b) {
  return a >= b ? (int64_t)a * (a + 1) / 2 + b : (int64_t)b * (b + 1) / 2 + a;
}

static void lt_pair_lk(int64_t idx, int64_t *a, int64_t *b) {
  int64_t r = (int64_t)((sqrt(8.0 * (double)idx + 1.0) - 1.0) * 0.5);
  while (r * (r + 1) / 2 > idx) --r;
  while ((r + 1) * (r + 2) / 2 <= idx) ++r;
  *a = r;
  *b = idx - r * (r + 1) / 2;
}

/* Ylk, MLk and the sorted, cut shell list of orbital k of density d. */
static void lk_screen_orbital(lk_state *s, int32_t d, int32_t k) {
  const int32_t n = s->n, ns = s->nshell;
  const double *c = s->csq[d] + (int64_t)n * k;
  for (int32_t i = 0; i < n; ++i) s->absc[i] = fabs(c[i]);
  /* Y(l)[k] = sum_n DH(l,n)|C(n)[k]| (DGEMV 'N', :584-590) */
  for (int32_t l = 0; l < n; ++l) {
    double y = 0.0;
    for (int32_t m = 0; m < n; ++m) y += s->diah[(int64_t)m * n + l] * s->absc[m];
    s->ylk[l] = y;
  }
  double *mlk = s->mlk + ((int64_t)d * s->maxocc + k) * ns;
  int32_t *indx = s->indx + ((int64_t)d * s->maxocc + k) * (ns + 1);
  for (int32_t sh = 0; sh < ns; ++sh) {
    double ym = 0.0;
    for (int32_t i = 0; i < s->nbassh[sh]; ++i) {
      const double y = s->ylk[s->koff[sh] + i];
      if (y > ym) ym = y; /* max(YshMax, Y) */
    }
    mlk[sh] = ym;
    indx[1 + sh] = sh;
  }
  /* selection sort with the cut, :602-627 / cho_lk_casscf.F90:656-698 */
  int32_t numsh = 0;
  for (int32_t jml = 0; jml < ns; ++jml) {
    double ymax = mlk[jml];
    int32_t jmax = jml;
    for (int32_t iml = jml + 1; iml < ns; ++iml)
      if (mlk[iml] > ymax) {
        ymax = mlk[iml];
        jmax = iml;
      }
    if (jmax != jml) {
      const double xt = mlk[jml];
      const int32_t it = indx[1 + jml];
      mlk[jml] = ymax;
      indx[1 + jml] = indx[1 + jmax];
      mlk[jmax] = xt;
      indx[1 + jmax] = it;
    }
    const int keep = s->cas_cut ? (mlk[jml] >= s->xtau[d]) : (mlk[jml] * mlk[0] >= s->tau[d]);
    if (!keep) break;
    ++numsh;
  }
  indx[0] = numsh;
  if (s->stats) s->stats[0] += numsh;
}

/* One batch of jnum vectors lrs(nrs, jnum) on reduced set `map`. */
static void lk_batch(lk_state *s, const int64_t *map, int64_t nrs, int64_t jnum,
                     const double *lrs, int doscreen, double *klt) {
  const int32_t n = s->n, ns = s->nshell;
  const int64_t np = (int64_t)n * (n + 1) / 2, nshp = (int64_t)ns * (ns + 1) / 2;
  if (s->estimate) { /* :461-479: D(ab) = sum_J L^2 of this batch */
    memset(s->diag, 0, sizeof(double) * np);
    for (int64_t r = 0; r < nrs; ++r)
      for (int64_t j = 0; j < jnum; ++j) s->diag[map[r]] += lrs[r + j * nrs] * lrs[r + j * nrs];
  }
  /* CHO_GetShFull: full square vectors and SvShp = max_J sum_{ab in shp} L^2 */
  memset(s->svshp, 0, sizeof(double) * nshp);
  memset(s->sv2, 0, sizeof(double) * nshp);
  memset(s->nrs_shp, 0, sizeof(int64_t) * nshp);
  memset(s->lfull, 0, sizeof(double) * jnum * n * n);
  for (int64_t r = 0; r < nrs; ++r) {
    int64_t a, b;
    lt_pair_lk(map[r], &a, &b);
    s->nrs_shp[shp_index(s->shell[a], s->shell[b])] += 1;
  }
  for (int64_t j = 0; j < jnum; ++j) {
    double *lf = s->lfull + j * n * n;
    for (int64_t r = 0; r < nrs; ++r) {
      int64_t a, b;
      lt_pair_lk(map[r], &a, &b);
      const double v = lrs[r + j * nrs];
      lf[a * n + b] = v;
      lf[b * n + a] = v;
      s->sv2[shp_index(s->shell[a], s->shell[b])] += v * v;
    }
    for (int64_t p = 0; p < nshp; ++p) {
      if (s->sv2[p] > s->svshp[p]) s->svshp[p] = s->sv2[p]; /* max(SvShp(:,1), SvShp(:,2)) */
      s->sv2[p] = 0.0;
    }
  }
  if (doscreen) /* swap_tosqrt (src/fock_util/swap_tosqrt.F90:57-72) */
    for (int64_t a = 0; a < n; ++a)
      for (int64_t b = 0; b <= a; ++b) {
        const double v = sqrt(fabs(s->diag[a * (a + 1) / 2 + b]));
        s->diah[a * n + b] = v;
        s->diah[b * n + a] = v;
      }
  for (int32_t d = 0; d < s->nden; ++d) {
    double *kd = klt + (int64_t)d * np;
    for (int32_t k = 0; k < s->nocc[d]; ++k) {
      if (doscreen) lk_screen_orbital(s, d, k);
      const double *ck = s->csq[d] + (int64_t)n * k;
      const double *sac = s->sumac + ((int64_t)d * s->maxocc + k) * ns;
      const double *mlk = s->mlk + ((int64_t)d * s->maxocc + k) * ns;
      const int32_t *indx = s->indx + ((int64_t)d * s->maxocc + k) * (ns + 1);
      const int32_t numsh = indx[0];
      memset(s->lab, 0, sizeof(double) * jnum * n);
      /* half-transform of the listed shells, :659-702 */
      for (int32_t ish = 0; ish < numsh; ++ish) {
        const int32_t sa = indx[1 + ish];
        s->keep[sa] = 1;
        int32_t ibcount = 0;
        for (int32_t sb = 0; sb < ns; ++sb) {
          const int64_t p = shp_index(sa, sb);
          if (s->nrs_shp[p] * s->nbassh[sa] * s->nbassh[sb] <= 0) continue;
          if (!(sqrt(fabs(sac[sb] * s->svshp[p])) >= s->thrv[d])) continue;
          ++ibcount;
          if (s->stats) s->stats[1] += 1;
          /* L(aJ)[k] += sum_{b in sb} L(aJ,b) C(b)[k] (DGEMV 'N', column by column) */
          for (int64_t bl = 0; bl < s->nbassh[sb]; ++bl) {
            const int64_t bg = s->koff[sb] + bl;
            const double cb = ck[bg];
            for (int64_t j = 0; j < jnum; ++j)
              for (int64_t al = 0; al < s->nbassh[sa]; ++al) {
                const int64_t ag = s->koff[sa] + al;
                s->lab[j * n + ag] += s->lfull[j * n * n + ag * n + bg] * cb;
              }
          }
        }
        if (ibcount == 0) s->keep[sa] = 0;
      }
      /* Faa, :713-743 */
      for (int32_t ish = 0; ish < numsh; ++ish) {
        const int32_t sa = indx[1 + ish];
        if (!s->keep[sa]) continue;
        double t = 0.0;
        for (int64_t al = 0; al < s->nbassh[sa]; ++al) {
          const int64_t ag = s->koff[sa] + al;
          double f = 0.0;
          for (int64_t j = 0; j < jnum; ++j) f += s->lab[j * n + ag] * s->lab[j * n + ag];
          if (fabs(f) > t) t = fabs(f);
        }
        s->faa[sa] = t;
      }
      /* K blocks of the interacting listed shell pairs, :749-830 */
      for (int32_t lsh = 0; lsh < numsh; ++lsh) {
        const int32_t sa = indx[1 + lsh];
        for (int32_t msh = 0; msh < numsh; ++msh) {
          const int32_t sb = indx[1 + msh];
          if (mlk[lsh] * mlk[msh] < s->tau[d]) break; /* the rest of the list is smaller */
          const double xfab = sqrt(fabs(s->faa[sa] * s->faa[sb]));
          const int diag_blk = sa == sb && xfab >= s->tau[d] / s->maxredt && s->keep[sa];
          const int off_blk = sa > sb && xfab >= s->tau[d] / s->maxredt && s->keep[sa] && s->keep[sb];
          if (!diag_blk && !off_blk) continue;
          if (s->stats) s->stats[2] += 1;
          const int64_t off = s->poff[shp_index(sa, sb)];
          for (int64_t bl = 0; bl < s->nbassh[sb]; ++bl)
            for (int64_t al = (sa == sb ? bl : 0); al < s->nbassh[sa]; ++al) {
              const int64_t ag = s->koff[sa] + al, bg = s->koff[sb] + bl;
              double acc = 0.0;
              for (int64_t j = 0; j < jnum; ++j) acc += s->lab[j * n + ag] * s->lab[j * n + bg];
              const int64_t idx = sa == sb ? al * (al + 1) / 2 + bl : s->nbassh[sa] * bl + al;
              kd[off + idx] += s->fact[d] * acc;
            }
        }
      }
    }
  }
  if (s->update && s->nscreen > 0) /* :834-880: D(ab) -= sum_J L^2, vector by vector */
    for (int64_t r = 0; r < nrs; ++r)
      for (int64_t j = 0; j < jnum; ++j) s->diag[map[r]] = s->diag[map[r]] - lrs[r + j * nrs] * lrs[r + j * nrs];
}

/* The LK exchange of nden densities with orbitals csq[d] (square A2(n,n), AO
 * index fastest, first nocc[d] columns): klt[d] (shell-pair-blocked LT,
 * nbas(nbas+1)/2 each) += fact[d] * the screened sum_{J,k} X X^T.  tau/thrv
 * as ora_lk_thresholds; cas_cut selects the CASSCF list cut.  stats (or
 * NULL) receives [listed (batch, orbital, shell), half-transformed shell
 * pairs, K blocks].  Returns 0, or 99 on bad shells. */
int32_t ora_lk_exchange(const ora_chol *ch, int32_t nden, const int32_t *nocc,
                        const double *const *csq, const double *fact, const double *tau,
                        const double *thrv, int32_t cas_cut, const ora_lk *lk,
                        int64_t nvec_batch, double *klt, int64_t *stats) {
  const int32_t n = ch->n, ns = lk->nshell;
  const int64_t np = (int64_t)n * (n + 1) / 2, nshp = (int64_t)ns * (ns + 1) / 2;
  int64_t tot = 0;
  for (int32_t s = 0; s < ns; ++s) tot += lk->nbassh[s];
  if (ns < 1 || tot != n || nden < 1 || nden > 2) return 99;
  lk_state s;
  memset(&s, 0, sizeof(s));
  s.n = n;
  s.nshell = ns;
  s.nbassh = lk->nbassh;
  s.nden = nden;
  s.cas_cut = cas_cut;
  s.maxredt = (double)lk->maxred;
  s.estimate = lk->estimate;
  s.update = lk->update;
  s.nscreen = lk->nscreen;
  s.stats = stats;
  if (stats) memset(stats, 0, 3 * sizeof(int64_t));
  s.maxocc = 1;
  for (int32_t d = 0; d < nden; ++d) {
    s.nocc[d] = nocc[d];
    s.csq[d] = csq[d];
    s.fact[d] = fact[d];
    s.tau[d] = tau[d];
    s.thrv[d] = thrv[d];
    s.xtau[d] = sqrt(tau[d]);
    if (nocc[d] > s.maxocc) s.maxocc = nocc[d];
  }
  s.koff = malloc(sizeof(int64_t) * ns);
  s.shell = malloc(sizeof(int32_t) * n);
  s.poff = malloc(sizeof(int64_t) * nshp);
  int64_t ao = 0, lk_off = 0;
  for (int32_t a = 0; a < ns; ++a) {
    s.koff[a] = ao;
    for (int32_t i = 0; i < lk->nbassh[a]; ++i) s.shell[ao + i] = a;
    ao += lk->nbassh[a];
    for (int32_t b = 0; b <= a; ++b) { /* nnBfShp, :276-292 */
      s.poff[(int64_t)a * (a + 1) / 2 + b] = lk_off;
      lk_off += (int64_t)lk->nbassh[a] * lk->nbassh[b] -
                (a == b ? (int64_t)lk->nbassh[a] * (lk->nbassh[a] - 1) / 2 : 0);
    }
  }
  s.diag = malloc(sizeof(double) * np);
  if (lk->diag) memcpy(s.diag, lk->diag, sizeof(double) * np); /* CHO_IODIAG read, :180 */
  else memset(s.diag, 0, sizeof(double) * np);
  s.diah = calloc((size_t)n * n, sizeof(double));
  s.sumac = calloc((size_t)nden * s.maxocc * ns, sizeof(double));
  s.mlk = calloc((size_t)nden * s.maxocc * ns, sizeof(double));
  s.indx = calloc((size_t)nden * s.maxocc * (ns + 1), sizeof(int32_t));
  s.svshp = malloc(sizeof(double) * nshp);
  s.sv2 = malloc(sizeof(double) * nshp);
  s.nrs_shp = malloc(sizeof(int64_t) * nshp);
  s.ylk = malloc(sizeof(double) * n);
  s.absc = malloc(sizeof(double) * n);
  s.faa = calloc((size_t)ns, sizeof(double));
  s.keep = calloc((size_t)ns, 1);
  /* SumAClk = sum_{i in shell} C(i,k)^2, :235-253 */
  for (int32_t d = 0; d < nden; ++d)
    for (int32_t k = 0; k < nocc[d]; ++k)
      for (int32_t sh = 0; sh < ns; ++sh) {
        double t = 0.0;
        for (int32_t i = 0; i < lk->nbassh[sh]; ++i) {
          const double c = csq[d][(int64_t)n * k + s.koff[sh] + i];
          t += c * c;
        }
        s.sumac[((int64_t)d * s.maxocc + k) * ns + sh] = t;
      }
  int32_t rc = 0;
  if (ch->nvec > 0) {
    const int32_t jred1 = ch->vred[0], jred2 = ch->vred[ch->nvec - 1];
    int64_t maxnrs = 1;
    for (int32_t r = 0; r < ch->nred; ++r)
      if (ch->nrs[r] > maxnrs) maxnrs = ch->nrs[r];
    const int64_t nb = nvec_batch < ch->nvec ? nvec_batch : ch->nvec;
    double *lrs = malloc(sizeof(double) * maxnrs * nb);
    s.lfull = malloc(sizeof(double) * nb * n * n);
    s.lab = malloc(sizeof(double) * nb * n);
    int32_t kscreen = 1, doscreen = 1; /* :336-337 */
    for (int32_t jred = jred1; jred <= jred2; ++jred) {
      int64_t ivrs = -1, nvrs = 0; /* Cho_X_nVecRS */
      for (int64_t j = 0; j < ch->nvec; ++j)
        if (ch->vred[j] == jred) {
          if (ivrs < 0) ivrs = j;
          ++nvrs;
        }
      if (nvrs > 0) {
        const int64_t nrs = ch->nrs[jred];
        const int64_t *map = ch->rs2lt + ch->rs_off[jred];
        const int64_t nvec = nvec_batch < nvrs ? nvec_batch : nvrs; /* :374 */
        const int64_t nbatch = (nvrs - 1) / nvec + 1;
        for (int64_t ib = 0; ib < nbatch; ++ib) {
          const int64_t jnum = ib == nbatch - 1 ? nvrs - nvec * (nbatch - 1) : nvec;
          const int64_t jvec = nvec * ib + ivrs;
          for (int64_t j = 0; j < jnum; ++j)
            memcpy(lrs + j * nrs, ch->lvec + ch->voff[jvec + j], sizeof(double) * nrs);
          lk_batch(&s, map, nrs, jnum, lrs, doscreen, klt);
          doscreen = 0; /* :831 */
        }
      }
      /* screening control, :891-898 */
      doscreen = kscreen == lk->nscreen;
      if (!doscreen) ++kscreen;
      else kscreen = 1;
    }
    free(lrs);
  }
  free(s.koff); free(s.shell); free(s.poff); free(s.diag); free(s.diah); free(s.sumac);
  free(s.mlk); free(s.indx); free(s.svshp); free(s.sv2); free(s.nrs_shp); free(s.lfull);
  free(s.lab); free(s.ylk); free(s.absc); free(s.faa); free(s.keep);
  return rc;
}

/* KLT (shell-pair blocked) -> FLT (LT), cho_lk_scf.F90:963-1028. */
static void lk_fold(int32_t n, const ora_lk *lk, const double *kd, double *fd) {
  const int32_t ns = lk->nshell;
  int64_t ao_a = 0, off = 0;
  for (int32_t sa = 0; sa < ns; ++sa) {
    int64_t ao_b = 0;
    for (int32_t sb = 0; sb <= sa; ++sb) {
      for (int64_t bl = 0; bl < lk->nbassh[sb]; ++bl)
        for (int64_t al = (sa == sb ? bl : 0); al < lk->nbassh[sa]; ++al) {
          const int64_t ag = ao_a + al, bg = ao_b + bl;
          const int64_t idx = sa == sb ? al * (al + 1) / 2 + bl : lk->nbassh[sa] * bl + al;
          fd[ag * (ag + 1) / 2 + bg] += kd[off + idx];
        }
      off += sa == sb ? (int64_t)lk->nbassh[sa] * (lk->nbassh[sa] + 1) / 2
                      : (int64_t)lk->nbassh[sa] * lk->nbassh[sb];
      ao_b += lk->nbassh[sb];
    }
    ao_a += lk->nbassh[sa];
  }
  (void)n;
}

/* CHO_LK_SCF (src/fock_util/cho_lk_scf.F90:14-1103), nSym = 1, screened:
 * FLT(d) += J[PLT] + KLT(d) with KLT(d) += -FactXI * (screened K).  porb[d]
 * is the square A2(n,n) (AO index fastest), first nforb+niorb columns.  klt
 * (or NULL) receives the shell-pair-blocked KLT. */
int32_t ora_cho_lk_scf_scr(const ora_chol *ch, int32_t nden, double *flt, double *klt,
                           const int32_t *nforb, const int32_t *niorb,
                           const double *const *porb, const double *plt, double factxi,
                           int32_t nscreen, double dmpk, double dfmat, const ora_lk *lk,
                           int64_t nvec_batch, int64_t *stats) {
  if (nden != 1 && nden != 2) return 99; /* :93-96 */
  const int32_t n = ch->n;
  const int64_t np = (int64_t)n * (n + 1) / 2;
  int32_t nocc[2] = {0, 0};
  for (int32_t d = 0; d < nden; ++d) nocc[d] = nforb[d] + niorb[d];
  /* Coulomb of PLT to every density (the exchange factor 0 leaves KLT zero) */
  int32_t zero[2] = {0, 0};
  int32_t rc = ora_cho_lk_scf(ch, nden, flt, NULL, zero, zero, porb, plt, 0.0, 1, &n, nvec_batch);
  if (rc) return rc;
  double tau[2], thrv[2];
  ora_lk_thresholds(lk->thrcom, dfmat, dmpk, nden, nocc, ch->nvec, lk->estimate, lk->maxred, tau, thrv);
  ora_lk nl = *lk;
  nl.nscreen = nscreen;
  double *kbuf = klt ? klt : calloc((size_t)nden * np, sizeof(double));
  double *kcur = calloc((size_t)nden * np, sizeof(double));
  const double fact[2] = {-factxi, -factxi}; /* :803,830 */
  rc = ora_lk_exchange(ch, nden, nocc, porb, fact, tau, thrv, 0, &nl, nvec_batch, kcur, stats);
  if (!rc)
    for (int32_t d = 0; d < nden; ++d) {
      for (int64_t p = 0; p < np; ++p) kbuf[d * np + p] += kcur[d * np + p];
      lk_fold(n, lk, kcur + d * np, flt + d * np);
    }
  free(kcur);
  if (!klt) free(kbuf);
  return rc;
}

/* CHO_LK_CASSCF (src/fock_util/cho_lk_casscf.F90:14-1268), nSym = 1,
 * screened: J of DI/DA and (pw|xy) as the unscreened routine, the exchange
 * K[MSQ] (FactXI*ExFac) and K[Ash(2)] (-1/2*ExFac, :113) screened with the
 * CASSCF list cut. */
int32_t ora_cho_lk_casscf_scr(const ora_chol *ch, const double *dlt, double *flt,
                              const double *msq, double *w_pwxy, double factxi,
                              int32_t nfiorb, int32_t naorb, int32_t nchm,
                              const double *ash1, const double *ash2, int32_t do_active,
                              int32_t nscreen, double dmpk, double dfmat, const double *cmo,
                              int32_t nfro, int32_t norb, double exfac, const ora_lk *lk,
                              int64_t nvec_batch, int64_t *stats) {
  const int32_t n = ch->n;
  const int64_t np = (int64_t)n * (n + 1) / 2;
  /* Coulomb and (pw|xy) with the exchange switched off (no orbitals) */
  int32_t rc = ora_cho_lk_casscf(ch, dlt, flt, msq, w_pwxy, factxi, 0, naorb, 0, ash1, ash2,
                                 do_active, cmo, nfro, norb, exfac, nvec_batch);
  if (rc) return rc;
  const int32_t nden = do_active ? 2 : 1;
  const int32_t nocc[2] = {nfiorb, do_active ? nchm : 0};
  double tau[2], thrv[2];
  ora_lk_thresholds(lk->thrcom, dfmat, dmpk, nden, nocc, ch->nvec, lk->estimate, lk->maxred, tau, thrv);
  ora_lk nl = *lk;
  nl.nscreen = nscreen;
  const double *csq[2] = {msq, ash2};
  const double fact[2] = {factxi * exfac, -0.5 * exfac};
  double *kcur = calloc((size_t)nden * np, sizeof(double));
  rc = ora_lk_exchange(ch, nden, nocc, csq, fact, tau, thrv, 1, &nl, nvec_batch, kcur, stats);
  if (!rc)
    for (int32_t d = 0; d < nden; ++d) lk_fold(n, lk, kcur + d * np, flt + d * np);
  free(kcur);
  return rc;
}
