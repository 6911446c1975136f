// Cho_X_Init from the reference's own files: the restart file CHORST, the
// reduced-set index file CHRED, the diagonal file CHODIAG and the vector
// files CHVEC<iSym> (nSym = 1..8).  Host code: parsing needs no device; the
// loader feeds a context through the ordinary entry points.
//
// Layouts (all word-addressed DA files, 8-byte words, 8-byte integers):
//   CHORST (src/cholesky_util/cho_wrrstc.F90:30-92, read back by
//     cho_x_rdrst.F90:33-190): nSym, nShell, nnShl, 0 | nBas(1:nSym) |
//     ScDiag, Cho_AdrVec | ThrCom, ThrDiag, Damp(1:2), Span, ThrNeg,
//     WarNeg, TooNeg (doubles) | MaxRed | InfRed(1:MaxRed) | per irrep:
//     NumCho, then InfVec(1:NumCho, j) for j = 1..5 (cho_wrrstc.F90:62-74).
//   CHRED (cho_getred.F90:46-78): reduced set IPASS at word InfRed(IPASS):
//     nnBstRSh(1:nSym, 1:nnShl) (irrep fastest) | IndRed(1:sum nnBstRSh) |
//     for IPASS = 1 also IndRSh(same length) | iSP2F(1:nnShl).  The elements
//     run irrep by irrep, shell pair by shell pair inside an irrep
//     (cho_setredind.F90: iiBstR, iiBstRSh).
//     For IPASS = 1, IndRed holds the shell-pair-local index LAB of each
//     element; for IPASS > 1 its position in reduced set 1.
//   CHODIAG (cho_iodiag.F90:28-33): the diagonal, nnBstRT(1) doubles in the
//     order of reduced set 1, from byte 0.
// The AO pair of a reduced-set-1 element is rebuilt as CHO_RSTOF does it
// (cho_rstof.F90:25-89): shell pair (ShlA >= ShlB) from iSP2F, LAB split into
// the local indices (iTri for ShlA = ShlB, NBstSh(ShlA)*(LB-1)+LA otherwise),
// local index -> SO with iShlSO (cho_setsh2.F90:22-33) from the caller's
// iSOShl (the runfile array 'ISOSHL'; SOs numbered irrep by irrep).  With
// nSym > 1 the decoded pair (a, b) must lie in the irrep block it is filed
// under (irrep(a) x irrep(b) = iSym, cho_rstof.F90:34-60), and every
// (irrep, reduced set) becomes a map of its own into the C1 triangle of all
// nBasT SOs, numbered (iSym-1)*MaxRed + IRED-1.
#include <fcntl.h>
#include <unistd.h>

#include <cerrno>
#include <cmath>
#include <cstring>

#include "mjk_internal.h"

struct mjk_choinfo {
  int32_t nsym = 0, nshell = 0, nnshl = 0, nbas = 0, adrvec = 0, maxred = 0;
  int32_t nbas_sym[8] = {};
  int64_t numcho_sym[8] = {};
  double thrcom = 0.0;
  int64_t numcho = 0;                            // all irreps
  std::vector<int64_t> infred;
  std::vector<int64_t> infvec[5];                // InfVec(J, 1..5, iSym), irrep 1's vectors first;
                                                 // column 2 holds the map number (iSym-1)*MaxRed + IRED
  std::vector<int32_t> vsym;                     // 0-based irrep of each vector
  std::vector<std::vector<int64_t>> rs2lt;       // per (irrep, reduced set): LT index of each element
  std::vector<int64_t> set1;                     // reduced set 1, all irreps, in file order
  std::vector<int32_t> nbassh;                   // shell sizes when shells are contiguous
  std::string err;
};

namespace {

struct WordFile {
  int fd = -1;
  ~WordFile() {
    if (fd >= 0) close(fd);
  }
  bool open_ro(const char *path) {
    fd = open(path, O_RDONLY);
    return fd >= 0;
  }
  // n 8-byte words at word address adr
  bool read(int64_t adr, int64_t n, void *dst) const {
    char *p = static_cast<char *>(dst);
    int64_t left = n * 8, at = adr * 8;
    while (left > 0) {
      const ssize_t got = pread(fd, p, size_t(left), off_t(at));
      if (got < 0 && errno == EINTR) continue;
      if (got <= 0) return false;
      p += got;
      left -= got;
      at += got;
    }
    return true;
  }
};

int fail(mjk_choinfo *ci, int rc, const std::string &msg) {
  ci->err = msg;
  return rc;
}

// 1-based iTri inverse: ij -> (i, j), i >= j (iTri_Rev).
void itri_rev(int64_t ij, int64_t *i, int64_t *j) {
  int64_t r = int64_t((std::sqrt(8.0 * double(ij) - 7.0) + 1.0) * 0.5);
  while (r * (r - 1) / 2 >= ij) --r;
  while ((r + 1) * r / 2 < ij) ++r;
  *i = r;
  *j = ij - r * (r - 1) / 2;
}

}  // namespace

extern "C" {

int mjk_choinfo_read(mjk_choinfo **out, const char *chorst, const char *chred,
                     const int32_t *isoshl) {
  if (!out) return MJK_ERR_ARG;
  *out = nullptr;
  auto *ci = new (std::nothrow) mjk_choinfo();
  if (!ci) return MJK_ERR_NOMEM;
  *out = ci;
  if (!chorst || !chred || !isoshl) return fail(ci, MJK_ERR_ARG, "choinfo_read: null argument");
  WordFile rst, red;
  if (!rst.open_ro(chorst))
    return fail(ci, MJK_ERR_IO, std::string("cannot open ") + chorst + ": " + std::strerror(errno));
  if (!red.open_ro(chred))
    return fail(ci, MJK_ERR_IO, std::string("cannot open ") + chred + ": " + std::strerror(errno));
  // ---- CHORST
  int64_t w[8];
  int64_t adr = 0;
  if (!rst.read(adr, 4, w)) return fail(ci, MJK_ERR_IO, "CHORST: header");
  adr += 4;
  if (w[0] < 1 || w[0] > 8) return fail(ci, MJK_ERR_ARG, "CHORST: nSym outside 1..8");
  const int32_t nsym = int32_t(w[0]);
  ci->nsym = nsym;
  ci->nshell = int32_t(w[1]);
  ci->nnshl = int32_t(w[2]);
  if (ci->nshell < 1 || ci->nnshl < 1 || ci->nnshl > int64_t(ci->nshell) * (ci->nshell + 1) / 2)
    return fail(ci, MJK_ERR_ARG, "CHORST: bad shell counts");  // cho_x_rdrst.F90:42-56
  if (!rst.read(adr, nsym, w)) return fail(ci, MJK_ERR_IO, "CHORST: nBas");
  adr += nsym;
  int64_t nbast = 0;
  for (int32_t s = 0; s < nsym; ++s) {
    if (w[s] < 0) return fail(ci, MJK_ERR_ARG, "CHORST: nBas < 0");
    ci->nbas_sym[s] = int32_t(w[s]);
    nbast += w[s];
  }
  if (nbast < 1 || nbast > 46340) return fail(ci, MJK_ERR_ARG, "CHORST: nBasT outside 1..46340");
  ci->nbas = int32_t(nbast);
  if (!rst.read(adr, 2, w)) return fail(ci, MJK_ERR_IO, "CHORST: flags");
  adr += 2;
  ci->adrvec = int32_t(w[1]);
  if (ci->adrvec != 1 && ci->adrvec != 2)
    return fail(ci, MJK_ERR_ARG, "CHORST: vector address mode not 1 or 2");  // :86-92
  double dscr[8];
  if (!rst.read(adr, 8, dscr)) return fail(ci, MJK_ERR_IO, "CHORST: thresholds");
  adr += 8;
  ci->thrcom = dscr[0];
  if (!rst.read(adr, 1, w)) return fail(ci, MJK_ERR_IO, "CHORST: MaxRed");
  adr += 1;
  ci->maxred = int32_t(w[0]);
  if (ci->maxred < 1) return fail(ci, MJK_ERR_ARG, "CHORST: MaxRed < 1");
  ci->infred.resize(size_t(ci->maxred));
  if (!rst.read(adr, ci->maxred, ci->infred.data())) return fail(ci, MJK_ERR_IO, "CHORST: InfRed");
  adr += ci->maxred;
  if (ci->infred[0] != 0) return fail(ci, MJK_ERR_ARG, "CHORST: InfRed(1) != 0");  // :151-156
  // per irrep: NumCho(iSym), then InfVec(1:NumCho, j, iSym), j = 1..5 (:159-176)
  const int32_t maxred = ci->maxred;
  for (int32_t s = 0; s < nsym; ++s) {
    if (!rst.read(adr, 1, w)) return fail(ci, MJK_ERR_IO, "CHORST: NumCho");
    adr += 1;
    const int64_t nc = w[0];
    if (nc < 0) return fail(ci, MJK_ERR_ARG, "CHORST: NumCho < 0");
    ci->numcho_sym[s] = nc;
    ci->numcho += nc;
    for (int j = 0; j < 5; ++j) {
      const size_t at = ci->infvec[j].size();
      ci->infvec[j].resize(at + size_t(nc));
      if (nc && !rst.read(adr, nc, ci->infvec[j].data() + at)) return fail(ci, MJK_ERR_IO, "CHORST: InfVec");
      adr += nc;
    }
    ci->vsym.insert(ci->vsym.end(), size_t(nc), s);
  }
  // ---- shells of the SOs: NBstSh and iShlSO (cho_setsh2.F90), irrep of each SO
  const int32_t n = ci->nbas, ns = ci->nshell, nnshl = ci->nnshl;
  std::vector<int64_t> nbstsh(static_cast<size_t>(ns), 0), ishlso(static_cast<size_t>(n), 0);
  std::vector<std::vector<int32_t>> so_of(static_cast<size_t>(ns));
  std::vector<int32_t> so_sym(static_cast<size_t>(n));
  for (int32_t s = 0, i = 0; s < nsym; ++s)
    for (int32_t k = 0; k < ci->nbas_sym[s]; ++k) so_sym[size_t(i++)] = s;
  for (int32_t i = 0; i < n; ++i) {
    const int32_t s = isoshl[i];
    if (s < 1 || s > ns) return fail(ci, MJK_ERR_ARG, "iSOShl: shell index out of range");
    ishlso[size_t(i)] = ++nbstsh[size_t(s - 1)];
    so_of[size_t(s - 1)].push_back(i);
  }
  bool contiguous = true;
  for (int32_t i = 1; i < n; ++i) contiguous = contiguous && isoshl[i] >= isoshl[i - 1];
  if (contiguous) ci->nbassh.assign(nbstsh.begin(), nbstsh.end());
  // ---- CHRED, reduced set 1: nnBstRSh, IndRed (= LAB), IndRSh, iSP2F
  const int64_t np = int64_t(n) * (n + 1) / 2;
  const int64_t nhead = int64_t(nsym) * nnshl;
  std::vector<int64_t> nnsh(static_cast<size_t>(nhead));
  if (!red.read(0, nhead, nnsh.data())) return fail(ci, MJK_ERR_IO, "CHRED: nnBstRSh of set 1");
  int64_t n1 = 0;
  for (int64_t v : nnsh) {
    if (v < 0) return fail(ci, MJK_ERR_ARG, "CHRED: negative nnBstRSh");
    n1 += v;
  }
  if (n1 > np) return fail(ci, MJK_ERR_ARG, "CHRED: reduced set 1 larger than the LT");
  std::vector<int64_t> lab(static_cast<size_t>(n1)), isp2f(static_cast<size_t>(nnshl));
  if ((n1 && !red.read(nhead, n1, lab.data())) || !red.read(nhead + 2 * n1, nnshl, isp2f.data()))
    return fail(ci, MJK_ERR_IO, "CHRED: index arrays of set 1");
  std::vector<int64_t> sa_of(static_cast<size_t>(nnshl)), sb_of(static_cast<size_t>(nnshl));
  for (int32_t sp = 0; sp < nnshl; ++sp) {  // full shell pair, sa >= sb (1-based)
    if (isp2f[size_t(sp)] < 1) return fail(ci, MJK_ERR_ARG, "CHRED: iSP2F out of range");
    itri_rev(isp2f[size_t(sp)], &sa_of[size_t(sp)], &sb_of[size_t(sp)]);
    if (sb_of[size_t(sp)] < 1 || sa_of[size_t(sp)] > ns) return fail(ci, MJK_ERR_ARG, "CHRED: iSP2F out of range");
  }
  std::vector<int64_t> &map1 = ci->set1;
  map1.resize(size_t(n1));
  std::vector<int32_t> sym1(static_cast<size_t>(n1));
  ci->rs2lt.assign(size_t(nsym) * maxred, {});
  int64_t k = 0;
  for (int32_t s = 0; s < nsym; ++s) {
    auto &ms = ci->rs2lt[size_t(s) * maxred];
    for (int32_t sp = 0; sp < nnshl; ++sp) {
      const int64_t sa = sa_of[size_t(sp)], sb = sb_of[size_t(sp)];
      for (int64_t e = 0; e < nnsh[size_t(s + int64_t(nsym) * sp)]; ++e, ++k) {
        const int64_t l = lab[size_t(k)];
        if (l < 1) return fail(ci, MJK_ERR_ARG, "CHRED: shell-pair index out of range");
        int64_t la, lb;
        if (sa == sb) {
          itri_rev(l, &la, &lb);
        } else {
          la = (l - 1) % nbstsh[size_t(sa - 1)] + 1;
          lb = (l - 1) / nbstsh[size_t(sa - 1)] + 1;
        }
        if (la > nbstsh[size_t(sa - 1)] || lb > nbstsh[size_t(sb - 1)])
          return fail(ci, MJK_ERR_ARG, "CHRED: shell-pair index out of range");
        const int64_t ia = so_of[size_t(sa - 1)][size_t(la - 1)];
        const int64_t ib = so_of[size_t(sb - 1)][size_t(lb - 1)];
        if ((so_sym[size_t(ia)] ^ so_sym[size_t(ib)]) != s)  // D2h product of 0-based irreps
          return fail(ci, MJK_ERR_ARG, "CHRED: element outside the irrep block it is filed under");
        const int64_t a = std::max(ia, ib), b = std::min(ia, ib);
        map1[size_t(k)] = a * (a + 1) / 2 + b;
        sym1[size_t(k)] = s;
        ms.push_back(map1[size_t(k)]);
      }
    }
  }
  // ---- reduced sets 2..MaxRed: IndRed = position in set 1
  for (int32_t r = 1; r < maxred; ++r) {
    const int64_t a0 = ci->infred[size_t(r)];
    if (a0 < 0) return fail(ci, MJK_ERR_ARG, "CHRED: negative reduced-set address");
    if (!red.read(a0, nhead, nnsh.data())) return fail(ci, MJK_ERR_IO, "CHRED: nnBstRSh");
    int64_t nr = 0;
    for (int64_t v : nnsh) {
      if (v < 0) return fail(ci, MJK_ERR_ARG, "CHRED: negative nnBstRSh");
      nr += v;
    }
    if (nr > n1) return fail(ci, MJK_ERR_ARG, "CHRED: reduced set larger than set 1");
    std::vector<int64_t> ind(static_cast<size_t>(nr));
    if (nr && !red.read(a0 + nhead, nr, ind.data())) return fail(ci, MJK_ERR_IO, "CHRED: IndRed");
    int64_t e = 0;
    for (int32_t s = 0; s < nsym; ++s) {
      auto &m = ci->rs2lt[size_t(s) * maxred + size_t(r)];
      int64_t cnt = 0;
      for (int32_t sp = 0; sp < nnshl; ++sp) cnt += nnsh[size_t(s + int64_t(nsym) * sp)];
      m.resize(size_t(cnt));
      for (int64_t q = 0; q < cnt; ++q, ++e) {
        const int64_t i1 = ind[size_t(e)];
        if (i1 < 1 || i1 > n1) return fail(ci, MJK_ERR_ARG, "CHRED: IndRed outside reduced set 1");
        if (sym1[size_t(i1 - 1)] != s)
          return fail(ci, MJK_ERR_ARG, "CHRED: element outside the irrep block it is filed under");
        m[size_t(q)] = map1[size_t(i1 - 1)];
      }
    }
  }
  // InfVec(J,2): reduced set, nondecreasing within an irrep; becomes the map number
  int64_t prev = 0;
  int32_t prev_sym = -1;
  for (int64_t j = 0; j < ci->numcho; ++j) {
    const int64_t r = ci->infvec[1][size_t(j)];
    const int32_t s = ci->vsym[size_t(j)];
    if (r < 1 || r > maxred) return fail(ci, MJK_ERR_ARG, "CHORST: InfVec(J,2) outside 1..MaxRed");
    if (s == prev_sym && r < prev) return fail(ci, MJK_ERR_ARG, "CHORST: vectors not ordered by reduced set");
    prev = r;
    prev_sym = s;
    ci->infvec[1][size_t(j)] = r + int64_t(s) * maxred;
  }
  return MJK_OK;
}

const char *mjk_choinfo_error(const mjk_choinfo *ci) { return ci ? ci->err.c_str() : "null"; }

int mjk_choinfo_free(mjk_choinfo *ci) {
  delete ci;
  return MJK_OK;
}

int mjk_choinfo_dims(const mjk_choinfo *ci, int32_t *nbas, int64_t *numcho, int32_t *maxred,
                     int32_t *nshell, int32_t *adrvec, double *thrcom) {
  if (!ci) return MJK_ERR_ARG;
  if (nbas) *nbas = ci->nbas;
  if (numcho) *numcho = ci->numcho;
  if (maxred) *maxred = ci->maxred;
  if (nshell) *nshell = ci->nshell;
  if (adrvec) *adrvec = ci->adrvec;
  if (thrcom) *thrcom = ci->thrcom;
  return MJK_OK;
}

int mjk_choinfo_symmetry(const mjk_choinfo *ci, int32_t *nsym, int32_t *nbas_sym, int64_t *numcho_sym) {
  if (!ci) return MJK_ERR_ARG;
  if (nsym) *nsym = ci->nsym;
  for (int s = 0; s < 8; ++s) {
    if (nbas_sym) nbas_sym[s] = ci->nbas_sym[s];
    if (numcho_sym) numcho_sym[s] = ci->numcho_sym[s];
  }
  return MJK_OK;
}

int mjk_choinfo_reduced_set(const mjk_choinfo *ci, int32_t jred, int64_t *nrs, int64_t *rs_to_lt) {
  if (!ci || !nrs || jred < 0 || ci->rs2lt.empty() ||
      int64_t(jred) >= int64_t(ci->nsym) * ci->maxred || ci->rs2lt.size() != size_t(ci->nsym) * ci->maxred)
    return MJK_ERR_ARG;
  const auto &m = ci->rs2lt[size_t(jred)];
  *nrs = int64_t(m.size());
  if (rs_to_lt) std::copy(m.begin(), m.end(), rs_to_lt);
  return MJK_OK;
}

int mjk_choinfo_infvec(const mjk_choinfo *ci, int32_t col, int64_t *values) {
  if (!ci || !values || col < 1 || col > 5) return MJK_ERR_ARG;
  std::copy(ci->infvec[col - 1].begin(), ci->infvec[col - 1].end(), values);
  return MJK_OK;
}

}  // extern "C"

namespace {

// The context from the files: every (irrep, reduced set) map declared, the
// vectors of irrep s read from chvec[s] (irrep 1's first), shells and the
// CHODIAG diagonal as mjk_load_choinfo describes.
int load_choinfo(mjk_context *c, const mjk_choinfo *ci, const char *const *chvec, int32_t nfiles,
                 const char *chodiag, int64_t split_bytes) {
  using namespace mjk;
  if (!c) return MJK_ERR_ARG;
  if (!ci || !chvec) return set_error(c, MJK_ERR_ARG, "load_choinfo: null argument");
  const size_t nmaps = size_t(ci->nsym) * size_t(ci->maxred);
  if (ci->maxred < 1 || ci->nsym < 1 || ci->rs2lt.size() != nmaps)
    return set_error(c, MJK_ERR_STATE, "load_choinfo: the index files were not read");
  if (nfiles != ci->nsym)
    return set_error(c, MJK_ERR_ARG, "load_choinfo: one CHVEC file per irrep is needed");
  for (int32_t s = 0; s < nfiles; ++s)
    if (!chvec[s] && ci->numcho_sym[s] > 0) return set_error(c, MJK_ERR_ARG, "load_choinfo: null CHVEC path");
  if (ci->nbas != c->n || ci->numcho > c->nchol)
    return set_error(c, MJK_ERR_ARG, "load_choinfo: context smaller than nBas / NumCho of the files");
  int rc;
  for (size_t r = 0; r < nmaps; ++r) {
    const auto &m = ci->rs2lt[r];
    if ((rc = mjk_set_reduced_set(c, int32_t(r), int64_t(m.size()), m.data()))) return rc;
  }
  // vectors map by map, at InfVec(J,3), MBL from Cho_AdrVec
  const int32_t mbl = ci->adrvec == 2 ? 512 : 8;
  for (int64_t j0 = 0; j0 < ci->numcho;) {
    const int64_t r = ci->infvec[1][size_t(j0)];
    const int32_t s = ci->vsym[size_t(j0)];
    int64_t j1 = j0;
    while (j1 < ci->numcho && ci->infvec[1][size_t(j1)] == r && ci->vsym[size_t(j1)] == s) ++j1;
    if ((rc = mjk_upload_vectors_file(c, chvec[s], int32_t(r - 1), j0, j1 - j0, ci->infvec[2].data() + j0, mbl,
                                      split_bytes)))
      return rc;
    j0 = j1;
  }
  if (!ci->nbassh.empty() && (rc = mjk_set_shells(c, ci->nshell, ci->nbassh.data()))) return rc;
  if (chodiag) {  // CHO_IODIAG: nnBstRT(1) doubles in reduced-set-1 order, all irreps
    const auto &m1 = ci->set1;
    std::vector<double> d1(m1.size()), dlt(static_cast<size_t>(c->np), 0.0);
    WordFile f;
    if (!f.open_ro(chodiag))
      return set_error(c, MJK_ERR_IO, std::string("load_choinfo: cannot open ") + chodiag);
    if (!m1.empty() && !f.read(0, int64_t(m1.size()), d1.data()))
      return set_error(c, MJK_ERR_SHORT_READ, "load_choinfo: CHODIAG shorter than reduced set 1");
    for (size_t e = 0; e < m1.size(); ++e) dlt[size_t(m1[e])] = d1[e];
    if ((rc = mjk_set_lk_diag(c, dlt.data(), ci->maxred, 0, 1))) return rc;
  }
  return MJK_OK;
}

}  // namespace

extern "C" {

int mjk_load_choinfo(mjk_context *c, const mjk_choinfo *ci, const char *chvec, const char *chodiag,
                     int64_t split_bytes) {
  if (c && ci && ci->nsym > 1)
    return mjk::set_error(c, MJK_ERR_ARG, "load_choinfo: nSym > 1, use mjk_load_choinfo_sym (one CHVEC per irrep)");
  return load_choinfo(c, ci, &chvec, 1, chodiag, split_bytes);
}

int mjk_load_choinfo_sym(mjk_context *c, const mjk_choinfo *ci, const char *const *chvec, int32_t nfiles,
                         const char *chodiag, int64_t split_bytes) {
  return load_choinfo(c, ci, chvec, nfiles, chodiag, split_bytes);
}

}  // extern "C"
