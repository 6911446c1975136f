// Pivoted incomplete Cholesky of a density (DECO), on the device.
//
// Reference: src/cholesky_util/cd_incore.F90:35-60 and
// src/cholesky_util/cd_incore_1.F90:14-105, used by src/scf/choscf_drv.F90:
// 99-135 (threshold 1e-8 * max diagonal) and src/fock_util/cho_cas_drv.F90:
// 129,227.  Each pass takes as pivot the diagonal element found by an
// in-order scan that only moves on when a value beats the running maximum by
// Acc = min(1e-12, Thr/100).  That pivot always lies within Acc of the largest
// diagonal, so a workgroup max-reduction over the diagonal (held in LDS) finds
// it whenever only one value is that close; near-ties replay the in-order scan
// on one lane.  The workgroup updates the pivot column, the new vector and the
// diagonal in parallel.  Contraction to FMA is off so that every update rounds
// like the reference's separate multiply and subtract.
#include "mjk_internal.h"

namespace mjk {
namespace {

constexpr int kCdThreads = 1024;
constexpr int kCdMaxN = 4096;

__device__ inline double block_max(double x, double *red) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) x = fmax(x, __shfl_xor(x, off, kWave));
  if (lane == 0) red[wave] = x;
  __syncthreads();
  double m = red[0];
  for (int w = 1; w < kCdThreads / kWave; ++w) m = fmax(m, red[w]);
  return m;
}

#pragma clang fp contract(off)
// X (n x n, column-major, overwritten), V (n x mxvec, column-major).
// state[0] = irc (0, 101, 102), state[1] = NumCho.
__global__ __launch_bounds__(kCdThreads) void k_cd_incore(double *__restrict__ X, int n,
                                                         double *__restrict__ V, int mxvec,
                                                         double thr, int32_t *__restrict__ state) {
  __shared__ double diag[kCdMaxN];
  __shared__ int32_t zlist[kCdMaxN];
  __shared__ double red[kCdThreads / kWave];
  __shared__ int32_t s_nz, s_fail, s_ncand, s_cand;
  __shared__ int32_t s_imax, s_stop;
  __shared__ double s_xmax;
  const double thr_fail = -1.0e-8, thr_neg = -1.0e-13;  // cd_incore.F90:50
  const double acc = fmin(1.0e-12, thr * 1.0e-2);        // cd_incore_1.F90:41
  const int tid = threadIdx.x;
  for (int i = tid; i < n; i += kCdThreads) diag[i] = X[int64_t(i) * n + i];
  int numcho = 0, irc = 0;
  for (int pass = 0; pass < n; ++pass) {
    if (tid == 0) {
      s_nz = 0;
      s_fail = 0;
      s_ncand = 0;
    }
    __syncthreads();
    // negative diagonals: fail below ThrFail, else zero row and column (:45-67)
    double m = 0.0;
    for (int i = tid; i < n; i += kCdThreads) {
      double d = diag[i];
      if (d < thr_neg) {
        if (d < thr_fail) {
          atomicOr(&s_fail, 1);
        } else {
          diag[i] = 0.0;
          d = 0.0;
          zlist[atomicAdd(&s_nz, 1)] = i;
        }
      }
      m = fmax(m, d);
    }
    m = block_max(m, red);  // includes a barrier
    // The pivot of the in-order scan (a value must beat the running maximum
    // by Acc, :68-71) lies within Acc of the largest diagonal m.  With one such
    // candidate it is the pivot; otherwise one lane replays the scan.
    for (int i = tid; i < n; i += kCdThreads)
      if (diag[i] >= m - acc) {
        atomicAdd(&s_ncand, 1);
        s_cand = i;
      }
    __syncthreads();
    if (tid == 0) {
      int stop = s_fail ? 101 : 0, imax = -1;
      double xmax = 0.0;
      if (!stop) {
        if (m <= acc) {
          xmax = 0.0;  // no value beats 0 + Acc
        } else if (s_ncand == 1) {
          imax = s_cand;
          xmax = diag[imax];
        } else {
          for (int i = 0; i < n; ++i)
            if (diag[i] > xmax + acc) {
              xmax = diag[i];
              imax = i;
            }
        }
        if (xmax <= thr) stop = -1;                      // converged (:74)
        else if (numcho == mxvec) stop = 102;            // too many vectors (:77-80)
      }
      s_imax = imax;
      s_xmax = xmax;
      s_stop = stop;
    }
    // zero the rows and columns of the negative diagonals
    const int nz = s_nz;
    for (int z = 0; z < nz; ++z) {
      const int64_t c = zlist[z];
      for (int i = tid; i < n; i += kCdThreads) {
        X[c * n + i] = 0.0;
        X[int64_t(i) * n + c] = 0.0;
      }
    }
    __syncthreads();
    const int stop = s_stop;
    if (stop) {
      irc = stop > 0 ? stop : 0;
      break;
    }
    const int imax = s_imax;
    const double xmax = s_xmax;
    const double xfac = 1.0 / sqrt(fabs(xmax));
    // X(:,imax) -= sum_j V(imax,j) V(:,j); X(imax,imax) = Xmax; new vector (:82-96)
    double *__restrict__ col = X + int64_t(imax) * n;
    double *__restrict__ vn = V + int64_t(numcho) * n;
    for (int i = tid; i < n; i += kCdThreads) {
      double x = col[i];
      for (int j = 0; j < numcho; ++j) x = x - V[int64_t(j) * n + imax] * V[int64_t(j) * n + i];
      if (i == imax) x = xmax;
      col[i] = x;
      const double d = (i == imax) ? xmax : diag[i];
      const double v = (d == 0.0) ? 0.0 : xfac * x;
      vn[i] = v;
      diag[i] = (i == imax) ? 0.0 : d - v * v;  // diagonal update (:98-101)
    }
    ++numcho;
    __syncthreads();
  }
  __syncthreads();
  // the destroyed X keeps the updated diagonal, as in the reference
  for (int i = tid; i < n; i += kCdThreads) X[int64_t(i) * n + i] = diag[i];
  if (tid == 0) {
    state[0] = irc;
    state[1] = numcho;
  }
}
#pragma clang fp contract(on)

}  // namespace
}  // namespace mjk

using namespace mjk;

extern "C" int mjk_cd_incore(mjk_context *c, int32_t n, double *x, double *vec, int32_t mxvec,
                             int32_t *numcho, double thr, int32_t *irc) {
  if (!c) return MJK_ERR_ARG;
  if (!numcho || !irc) return set_error(c, MJK_ERR_ARG, "cd_incore: null output");
  *numcho = 0;
  *irc = 0;
  if (n < 1) return MJK_OK;  // cd_incore.F90:53
  if (mxvec <= 0) {         // cd_incore.F90:58
    *irc = -1;
    return MJK_OK;
  }
  if (n > kCdMaxN) return set_error(c, MJK_ERR_ARG, "cd_incore: n > 4096 (one workgroup)");
  if (!x || !vec) return set_error(c, MJK_ERR_ARG, "cd_incore: null pointer");
  if (thr < 0.0) thr = 1.0e-6;  // DefThr, cd_incore.F90:50,56
  int rc;
  if ((rc = ensure(c, c->w->misc, 256))) return rc;
  int32_t *state = static_cast<int32_t *>(c->w->misc.ptr);
  hipLaunchKernelGGL(k_cd_incore, dim3(1), dim3(kCdThreads), 0, c->stream, x, int(n), vec,
                     int(mxvec), thr, state);
  if ((rc = hip_check(c, hipGetLastError(), "cd_incore launch"))) return rc;
  int32_t h[2] = {0, 0};
  if ((rc = hip_check(c, hipMemcpyAsync(h, state, sizeof(h), hipMemcpyDeviceToHost, c->stream),
                      "cd_incore state")) ||
      (rc = hip_check(c, hipStreamSynchronize(c->stream), "cd_incore sync")))
    return rc;
  *irc = h[0];
  *numcho = h[1];
  return MJK_OK;
}
