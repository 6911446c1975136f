// Cho_X_Init from the reference's own files: the restart file CHORST, the
// reduced-set index file CHRED, the diagonal file CHODIAG and the vector
// files CHVEC<iSym> (nSym = 1).  Host code: parsing needs no device; the
// loader feeds a context through the ordinary entry points.
//
// Layouts (all word-addressed DA files, 8-byte words, 8-byte integers):
//   CHORST (src/cholesky_util/cho_wrrstc.F90:30-92, read back by
//     cho_x_rdrst.F90:33-190): nSym, nShell, nnShl, 0 | nBas(1:nSym) |
//     ScDiag, Cho_AdrVec | ThrCom, ThrDiag, Damp(1:2), Span, ThrNeg,
//     WarNeg, TooNeg (doubles) | MaxRed | InfRed(1:MaxRed) | per irrep:
//     NumCho, then InfVec(1:NumCho, j) for j = 1..5.
//   CHRED (cho_getred.F90:46-78): reduced set IPASS at word InfRed(IPASS):
//     nnBstRSh(1:nSym, 1:nnShl) | IndRed(1:sum nnBstRSh) | for IPASS = 1
//     also IndRSh(same length) | iSP2F(1:nnShl).
//     For IPASS = 1, IndRed holds the shell-pair-local index LAB of each
//     element; for IPASS > 1 its position in reduced set 1.
//   CHODIAG (cho_iodiag.F90:28-33): the diagonal, nnBstRT(1) doubles in the
//     order of reduced set 1, from byte 0.
// The AO pair of a reduced-set-1 element is rebuilt as CHO_RSTOF does it
// (cho_rstof.F90:25-89): shell pair (ShlA >= ShlB) from iSP2F, LAB split into
// the local indices (iTri for ShlA = ShlB, NBstSh(ShlA)*(LB-1)+LA otherwise),
// local index -> SO with iShlSO (cho_setsh2.F90:22-33) from the caller's
// iSOShl (the runfile array 'ISOSHL').
#include <fcntl.h>
#include <unistd.h>

#include <cerrno>
#include <cmath>
#include <cstring>

#include "mjk_internal.h"

struct mjk_choinfo {
  int32_t nsym = 0, nshell = 0, nnshl = 0, nbas = 0, adrvec = 0, maxred = 0;
  double thrcom = 0.0;
  int64_t numcho = 0;
  std::vector<int64_t> infred;
  std::vector<int64_t> infvec[5];               // InfVec(J, 1..5)
  std::vector<std::vector<int64_t>> rs2lt;       // per reduced set: LT index of each element
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
  if (w[0] != 1) return fail(ci, MJK_ERR_ARG, "CHORST: nSym != 1 (only C1 is read)");
  ci->nsym = 1;
  ci->nshell = int32_t(w[1]);
  ci->nnshl = int32_t(w[2]);
  if (ci->nshell < 1 || ci->nnshl < 1 || ci->nnshl > int64_t(ci->nshell) * (ci->nshell + 1) / 2)
    return fail(ci, MJK_ERR_ARG, "CHORST: bad shell counts");  // cho_x_rdrst.F90:42-56
  if (!rst.read(adr, 1, w)) return fail(ci, MJK_ERR_IO, "CHORST: nBas");
  adr += 1;
  ci->nbas = int32_t(w[0]);
  if (ci->nbas < 1) return fail(ci, MJK_ERR_ARG, "CHORST: nBas < 1");
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
  if (!rst.read(adr, 1, w)) return fail(ci, MJK_ERR_IO, "CHORST: NumCho");
  adr += 1;
  ci->numcho = w[0];
  if (ci->numcho < 0) return fail(ci, MJK_ERR_ARG, "CHORST: NumCho < 0");
  for (int j = 0; j < 5; ++j) {
    ci->infvec[j].resize(size_t(ci->numcho));
    if (ci->numcho && !rst.read(adr, ci->numcho, ci->infvec[j].data()))
      return fail(ci, MJK_ERR_IO, "CHORST: InfVec");
    adr += ci->numcho;
  }
  // ---- shells of the SOs: NBstSh and iShlSO (cho_setsh2.F90)
  const int32_t n = ci->nbas, ns = ci->nshell;
  std::vector<int64_t> nbstsh(static_cast<size_t>(ns), 0), ishlso(static_cast<size_t>(n), 0);
  std::vector<std::vector<int32_t>> so_of(static_cast<size_t>(ns));
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
  std::vector<int64_t> nnsh(static_cast<size_t>(ci->nnshl));
  if (!red.read(0, ci->nnshl, nnsh.data())) return fail(ci, MJK_ERR_IO, "CHRED: nnBstRSh of set 1");
  int64_t n1 = 0;
  for (int64_t v : nnsh) {
    if (v < 0) return fail(ci, MJK_ERR_ARG, "CHRED: negative nnBstRSh");
    n1 += v;
  }
  if (n1 > np) return fail(ci, MJK_ERR_ARG, "CHRED: reduced set 1 larger than the LT");
  std::vector<int64_t> lab(static_cast<size_t>(n1)), isp2f(static_cast<size_t>(ci->nnshl));
  if ((n1 && !red.read(ci->nnshl, n1, lab.data())) ||
      !red.read(ci->nnshl + 2 * n1, ci->nnshl, isp2f.data()))
    return fail(ci, MJK_ERR_IO, "CHRED: index arrays of set 1");
  std::vector<int64_t> map1(static_cast<size_t>(n1));
  int64_t k = 0;
  for (int32_t sp = 0; sp < ci->nnshl; ++sp) {
    int64_t sa, sb;  // full shell pair, sa >= sb (1-based)
    if (isp2f[size_t(sp)] < 1) return fail(ci, MJK_ERR_ARG, "CHRED: iSP2F out of range");
    itri_rev(isp2f[size_t(sp)], &sa, &sb);
    if (sb < 1 || sa > ns) return fail(ci, MJK_ERR_ARG, "CHRED: iSP2F out of range");
    for (int64_t e = 0; e < nnsh[size_t(sp)]; ++e, ++k) {
      const int64_t l = lab[size_t(k)];
      if (l < 1) return fail(ci, MJK_ERR_ARG, "CHRED: shell-pair index out of range");
      int64_t la, lb;
      if (sa == sb) {
        itri_rev(l, &la, &lb);
      } else {
        la = (l - 1) % nbstsh[size_t(sa - 1)] + 1;
        lb = (l - 1) / nbstsh[size_t(sa - 1)] + 1;
      }
      if (l < 1 || la > nbstsh[size_t(sa - 1)] || lb > nbstsh[size_t(sb - 1)])
        return fail(ci, MJK_ERR_ARG, "CHRED: shell-pair index out of range");
      const int64_t ia = so_of[size_t(sa - 1)][size_t(la - 1)];
      const int64_t ib = so_of[size_t(sb - 1)][size_t(lb - 1)];
      const int64_t a = std::max(ia, ib), b = std::min(ia, ib);
      map1[size_t(k)] = a * (a + 1) / 2 + b;
    }
  }
  ci->rs2lt.resize(size_t(ci->maxred));
  ci->rs2lt[0] = map1;
  // ---- reduced sets 2..MaxRed: IndRed = position in set 1
  for (int32_t r = 1; r < ci->maxred; ++r) {
    const int64_t a0 = ci->infred[size_t(r)];
    if (a0 < 0) return fail(ci, MJK_ERR_ARG, "CHRED: negative reduced-set address");
    if (!red.read(a0, ci->nnshl, nnsh.data())) return fail(ci, MJK_ERR_IO, "CHRED: nnBstRSh");
    int64_t nr = 0;
    for (int64_t v : nnsh) nr += v;
    if (nr < 0 || nr > n1) return fail(ci, MJK_ERR_ARG, "CHRED: reduced set larger than set 1");
    std::vector<int64_t> ind(static_cast<size_t>(nr));
    if (nr && !red.read(a0 + ci->nnshl, nr, ind.data())) return fail(ci, MJK_ERR_IO, "CHRED: IndRed");
    auto &m = ci->rs2lt[size_t(r)];
    m.resize(size_t(nr));
    for (int64_t e = 0; e < nr; ++e) {
      if (ind[size_t(e)] < 1 || ind[size_t(e)] > n1)
        return fail(ci, MJK_ERR_ARG, "CHRED: IndRed outside reduced set 1");
      m[size_t(e)] = map1[size_t(ind[size_t(e)] - 1)];
    }
  }
  for (int64_t j = 0; j < ci->numcho; ++j) {
    const int64_t r = ci->infvec[1][size_t(j)];
    if (r < 1 || r > ci->maxred) return fail(ci, MJK_ERR_ARG, "CHORST: InfVec(J,2) outside 1..MaxRed");
    if (j && r < ci->infvec[1][size_t(j - 1)])
      return fail(ci, MJK_ERR_ARG, "CHORST: vectors not ordered by reduced set");
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

int mjk_choinfo_reduced_set(const mjk_choinfo *ci, int32_t jred, int64_t *nrs, int64_t *rs_to_lt) {
  if (!ci || !nrs || jred < 0 || jred >= ci->maxred || ci->rs2lt.size() != size_t(ci->maxred))
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

int mjk_load_choinfo(mjk_context *c, const mjk_choinfo *ci, const char *chvec, const char *chodiag,
                     int64_t split_bytes) {
  using namespace mjk;
  if (!c) return MJK_ERR_ARG;
  if (!ci || !chvec) return set_error(c, MJK_ERR_ARG, "load_choinfo: null argument");
  if (ci->rs2lt.size() != size_t(ci->maxred) || ci->maxred < 1)
    return set_error(c, MJK_ERR_STATE, "load_choinfo: the index files were not read");
  if (ci->nbas != c->n || ci->numcho > c->nchol)
    return set_error(c, MJK_ERR_ARG, "load_choinfo: context smaller than nBas / NumCho of the files");
  int rc;
  for (int32_t r = 0; r < ci->maxred; ++r) {
    const auto &m = ci->rs2lt[size_t(r)];
    if ((rc = mjk_set_reduced_set(c, r, int64_t(m.size()), m.data()))) return rc;
  }
  // vectors reduced set by reduced set, at InfVec(J,3), MBL from Cho_AdrVec
  const int32_t mbl = ci->adrvec == 2 ? 512 : 8;
  for (int64_t j0 = 0; j0 < ci->numcho;) {
    const int64_t r = ci->infvec[1][size_t(j0)];
    int64_t j1 = j0;
    while (j1 < ci->numcho && ci->infvec[1][size_t(j1)] == r) ++j1;
    if ((rc = mjk_upload_vectors_file(c, chvec, int32_t(r - 1), j0, j1 - j0,
                                      ci->infvec[2].data() + j0, mbl, split_bytes)))
      return rc;
    j0 = j1;
  }
  if (!ci->nbassh.empty() && (rc = mjk_set_shells(c, ci->nshell, ci->nbassh.data()))) return rc;
  if (chodiag) {  // CHO_IODIAG: nnBstRT(1) doubles in reduced-set-1 order
    const auto &m1 = ci->rs2lt[0];
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

}  // extern "C"
