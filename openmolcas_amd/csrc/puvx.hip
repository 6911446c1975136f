// (pw|xy) integrals of the CASSCF Fock build.
//
// Reference: src/fock_util/cho_fmcscf.F90:403-487 and
// src/fock_util/cho_eval_waxy.F90:12-150 (nSym = 1):
//   L(v,b,J)  = sum_a C(v,a) L_ab,J              (CHO_VTRA iSwap = 0)
//   Lxy(J)    = sum_b L(x,b,J) C(y,b), x <= y    (DGEMM_Tri 'N','T')
//   Scr(w,a,xy) += sum_J L(w,a,J) Lxy(J)          (DGEMM over the batch)
//   W(p,w,xy) = sum_a CMO(a,nfro+p) Scr(w,a,xy)   (after the last batch)
// The first step is the exchange half-transform with the active orbitals as
// the only bank set; the remaining contractions are small FP64 kernels.
#include "mjk_internal.h"

namespace mjk {
namespace {

// lxy[jr*nxy + xy] = sum_b Y[(jr*ldc + x)*n_pad + b] * C3[y + b*na], xy = y(y+1)/2 + x
__global__ void k_lxy(const double *__restrict__ Y, int ldc, int n, int n_pad,
                      const double *__restrict__ c3, int na, double *__restrict__ lxy) {
  const int64_t jr = blockIdx.x;
  const int nxy = na * (na + 1) / 2;
  for (int xy = threadIdx.x; xy < nxy; xy += blockDim.x) {
    int y = int((sqrt(8.0 * xy + 1.0) - 1.0) * 0.5);
    while (y * (y + 1) / 2 > xy) --y;
    while ((y + 1) * (y + 2) / 2 <= xy) ++y;
    const int x = xy - y * (y + 1) / 2;
    const double *__restrict__ yr = Y + (jr * ldc + x) * int64_t(n_pad);
    double s = 0.0;
    for (int b = 0; b < n; ++b) s += yr[b] * c3[y + int64_t(b) * na];
    lxy[jr * nxy + xy] = s;
  }
}

constexpr int kXB = 8;

// scr[(xy*na + w)*n_pad + a] += sum_jr Y[(jr*ldc + w)*n_pad + a] * lxy[jr*nxy + xy]
// grid = (ceil(na*n_pad/256), ceil(nxy/kXB))
__global__ void k_scr(const double *__restrict__ Y, int ldc, int n_pad, int na, int64_t jn,
                      const double *__restrict__ lxy, int nxy, double *__restrict__ scr) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= int64_t(na) * n_pad) return;
  const int w = int(i / n_pad), a = int(i % n_pad);
  const int xy0 = blockIdx.y * kXB;
  double acc[kXB] = {};
  for (int64_t jr = 0; jr < jn; ++jr) {
    const double yv = Y[(jr * ldc + w) * int64_t(n_pad) + a];
#pragma unroll
    for (int t = 0; t < kXB; ++t)
      if (xy0 + t < nxy) acc[t] += yv * lxy[jr * nxy + xy0 + t];
  }
#pragma unroll
  for (int t = 0; t < kXB; ++t)
    if (xy0 + t < nxy) scr[(int64_t(xy0 + t) * na + w) * n_pad + a] += acc[t];
}

// w[npw*xy + p + norb*w] = sum_a cmo[a + n*(nfro+p)] * scr[(xy*na + w)*n_pad + a]
__global__ void k_wpwxy(const double *__restrict__ scr, int na, int n, int n_pad,
                        const double *__restrict__ cmo, int nfro, int norb, int nxy,
                        double *__restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t npw = int64_t(norb) * na;
  if (i >= npw * nxy) return;
  const int64_t xy = i / npw;
  const int w = int((i % npw) / norb), p = int(i % norb);
  const double *__restrict__ cp = cmo + int64_t(n) * (nfro + p);
  const double *__restrict__ sr = scr + (xy * na + w) * int64_t(n_pad);
  double s = 0.0;
  for (int a = 0; a < n; ++a) s += cp[a] * sr[a];
  out[i] = s;
}

}  // namespace
}  // namespace mjk

using namespace mjk;

extern "C" int mjk_puvx(mjk_context *c, int32_t naorb, const double *porb3, const double *cmo,
                        int32_t nfro, int32_t norb, double *w) {
  if (!c) return MJK_ERR_ARG;
  if (naorb < 1 || naorb > 256 || !porb3 || !cmo || !w || nfro < 0 || norb < 0 ||
      nfro + norb > c->n)
    return set_error(c, MJK_ERR_ARG, "puvx: bad arguments");
  if (c->nloaded != c->nchol)
    return set_error(c, MJK_ERR_SHORT_READ, "puvx: not all Cholesky vectors were uploaded");
  const int n = c->n, n_pad = c->n_pad;
  const int nxy = naorb * (naorb + 1) / 2;
  const int64_t M = c->nchol;
  int rc;
  std::vector<int> off, opad;
  int ot = 0, tot = 0;
  const double *sets[1] = {porb3};
  if ((rc = pack_bank(c, 1, &naorb, sets, off, opad, &ot, &tot))) return rc;
  const int ldc = 16 * ot;
  const double *cb = static_cast<const double *>(c->w->cbank.ptr);
  const int64_t yvec = sizeof(double) * int64_t(ldc) * n_pad;
  int64_t jc = std::max<int64_t>(1, std::min<int64_t>(M, c->x_budget / yvec));
  const int64_t nchunk = ceil_div(M, jc);
  jc = ceil_div(M, nchunk);
  if ((rc = ensure(c, c->w->x, size_t(yvec * jc)))) return rc;
  double *Y = static_cast<double *>(c->w->x.ptr);
  const size_t scr_n = size_t(nxy) * naorb * n_pad;
  if ((rc = ensure(c, c->w->k, sizeof(double) * (scr_n + size_t(jc) * nxy)))) return rc;
  double *scr = static_cast<double *>(c->w->k.ptr);
  double *lxy = scr + scr_n;
  if ((rc = fill_zero(c, scr, sizeof(double) * scr_n))) return rc;
  for (int64_t ch = 0; ch < nchunk; ++ch) {
    const int64_t j0 = ch * jc, jn = std::min(jc, M - j0);
    phase_begin(c);
    ht_launch(c, ot, cb, j0, jn, Y);
    phase_end(c, MJK_PHASE_HALF_TRANSFORM);
    phase_begin(c);
    hipLaunchKernelGGL(k_lxy, dim3(unsigned(jn)), dim3(256), 0, c->stream, Y, ldc, n, n_pad, porb3,
                       naorb, lxy);
    hipLaunchKernelGGL(k_scr,
                       dim3(unsigned(ceil_div(int64_t(naorb) * n_pad, 256)),
                            unsigned(ceil_div(nxy, kXB))),
                       dim3(256), 0, c->stream, Y, ldc, n_pad, naorb, jn, lxy, nxy, scr);
    phase_end(c, MJK_PHASE_EXCHANGE);
  }
  phase_begin(c);
  const int64_t nout = int64_t(norb) * naorb * nxy;
  if (nout > 0)
    hipLaunchKernelGGL(k_wpwxy, dim3(unsigned(ceil_div(nout, 256))), dim3(256), 0, c->stream, scr,
                       naorb, n, n_pad, cmo, nfro, norb, nxy, w);
  phase_end(c, MJK_PHASE_ASSEMBLY);
  return hip_check(c, hipGetLastError(), "puvx launch");
}
