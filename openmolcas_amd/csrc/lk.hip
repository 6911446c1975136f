// Reference-form LK screening of the exchange: the per-orbital shell lists,
// vector-pair and shell-pair cuts of CHO_LK_SCF (src/fock_util/
// cho_lk_scf.F90:143-171,547-627,659-830) and CHO_LK_CASSCF (src/fock_util/
// cho_lk_casscf.F90:164-200,604-930), on the device.
//
// The reference decides, per batch of at most 25 vectors of one reduced set
// and per orbital k, which shells a are listed (Y = sqrt|D| |C_k|, its
// largest element per shell MLk, a cut on MLk), which shell pairs (a, b) enter
// the half-transform X_k = L C_k (sqrt|SumAClk(b,k) SvShp(ab)| >= thrv), and
// which K blocks (a >= b) receive X_k,a X_k,b^T (MLk(a) MLk(b) >= tau,
// sqrt|Faa(a) Faa(b)| >= tau/MaxRed, Keep flags).  The same decisions are
// made here; the arithmetic is tiled for the GPU:
//   - the half-transform stages 64 x 64 tiles of the packed vector in LDS and
//     each thread accumulates one AO row for four orbitals, the mask of a
//     (orbital, shell a, shell b) triple evaluated once per b-shell run;
//   - the K kernel owns a 64 x 64 tile of the lower triangle, stages the
//     batch's X rows of both tile sides per orbital and skips orbitals whose
//     mask is empty on the tile;
//   - decision quantities (diagonal, SvShp, SumAClk, Y, Faa) are evaluated
//     without FMA contraction, in the reference's summation order where it
//     has one, so the cuts fall where the reference's fall.
// The selection sort of the reference (:602-627) keeps the shells whose MLk
// passes the cut, in decreasing MLk order, stopping at the first failure;
// because the values along the sorted list do not increase and rounding is
// monotone, that set is exactly {s : MLk(s) max MLk >= tau} (SCF) or
// {s : MLk(s) >= sqrt(tau)} (CASSCF), which is what the list kernel forms.
#include <cmath>

#include "mjk_internal.h"

namespace mjk {
namespace {

constexpr int kLkBatch = 25;  // Cho_LK_MaxVecPerBatch (src/fock_util/cho_lk_maxvecperbatch.F90:30)
constexpr int kT = 64;        // AO tile
constexpr int kKc = 16;       // orbitals per half-transform workgroup
constexpr int kLkThreads = 256;

__device__ __forceinline__ int64_t shp_idx(int a, int b) {
  return a >= b ? int64_t(a) * (a + 1) / 2 + b : int64_t(b) * (b + 1) / 2 + a;
}

// One or two densities: orbital counts, first global orbital index, device
// orbital matrices (orbital-major), thresholds.
struct LkDens {
  int nd;
  int nocc[2];
  int k0[2];           // global orbital offset of density d
  int nkc[2];          // half-transform chunks of density d
  const double *c[2];  // C_d(b, k) at k + nocc[d]*b
  double tau[2], thrv[2], xtau[2];
  int ktot;            // sum of nocc
};

__device__ __forceinline__ void dk_of(const LkDens &D, int g, int *d, int *k) {
  *d = (D.nd > 1 && g >= D.k0[1]) ? 1 : 0;
  *k = g - D.k0[*d];
}

#pragma clang fp contract(off)

// Estimate: diag(p) = sum_J L_J(p)^2 of the batch (cho_lk_scf.F90:461-479).
__global__ void k_lk_est_diag(const double *__restrict__ L, int64_t ldl, int64_t j0, int jnum,
                              int64_t np, double *__restrict__ diag) {
  for (int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < np;
       p += int64_t(gridDim.x) * blockDim.x) {
    double s = 0.0;
    for (int j = 0; j < jnum; ++j) {
      const double v = L[(j0 + j) * ldl + p];
      s = s + v * v;
    }
    diag[p] = s;
  }
}

// Update: diag(p) = diag(p) - L_J(p)^2, vector by vector (:834-880).
__global__ void k_lk_update_diag(const double *__restrict__ L, int64_t ldl, int64_t j0, int jnum,
                                 int64_t np, double *__restrict__ diag) {
  for (int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < np;
       p += int64_t(gridDim.x) * blockDim.x) {
    double dv = diag[p];
    for (int j = 0; j < jnum; ++j) {
      const double v = L[(j0 + j) * ldl + p];
      dv = dv - v * v;
    }
    diag[p] = dv;
  }
}

// SvShp(ab) = max_J sum_{p in shell pair ab} L_J(p)^2 (CHO_GetShFull,
// src/fock_util/cho_getshfull.F90:48-96): one wave per shell pair.
__global__ __launch_bounds__(kLkThreads) void k_lk_svshp(const double *__restrict__ L, int64_t ldl,
                                                         int64_t j0, int jnum, int nshell,
                                                         const int32_t *__restrict__ koff,
                                                         const int32_t *__restrict__ nbsh,
                                                         double *__restrict__ sv) {
  const int lane = threadIdx.x & 63;
  const int64_t nshp = int64_t(nshell) * (nshell + 1) / 2;
  const int64_t p = int64_t(blockIdx.x) * (kLkThreads / 64) + (threadIdx.x >> 6);
  if (p >= nshp) return;
  int sa = int((sqrt(8.0 * double(p) + 1.0) - 1.0) * 0.5);
  while (int64_t(sa) * (sa + 1) / 2 > p) --sa;
  while (int64_t(sa + 1) * (sa + 2) / 2 <= p) ++sa;
  const int sb = int(p - int64_t(sa) * (sa + 1) / 2);
  const int na = nbsh[sa], nb = nbsh[sb], a0 = koff[sa], b0 = koff[sb];
  const int ne = sa == sb ? na * (na + 1) / 2 : na * nb;
  double best = 0.0;
  for (int j = 0; j < jnum; ++j) {
    const double *row = L + (j0 + j) * ldl;
    double s = 0.0;
    for (int e = lane; e < ne; e += 64) {
      int al, bl;
      if (sa == sb) {
        al = int((sqrt(8.0 * double(e) + 1.0) - 1.0) * 0.5);
        while (al * (al + 1) / 2 > e) --al;
        while ((al + 1) * (al + 2) / 2 <= e) ++al;
        bl = e - al * (al + 1) / 2;
      } else {
        al = e % na;
        bl = e / na;
      }
      const int64_t a = a0 + al, b = b0 + bl;
      const double v = row[a * (a + 1) / 2 + b];
      s = s + v * v;
    }
    for (int off = 32; off > 0; off >>= 1) s = s + __shfl_xor(s, off, 64);
    best = s > best ? s : best;  // max(SvShp(:,1), SvShp(:,2))
  }
  if (lane == 0) sv[p] = best;
}

// DH(a,b) = DH(b,a) = sqrt|D(ab)| (swap_tosqrt, src/fock_util/swap_tosqrt.F90:57-72).
__global__ void k_lk_diah(const double *__restrict__ diag, int n, double *__restrict__ dh) {
  const int a = blockIdx.x;
  for (int b = threadIdx.x; b <= a; b += blockDim.x) {
    const double v = sqrt(fabs(diag[int64_t(a) * (a + 1) / 2 + b]));
    dh[int64_t(a) * n + b] = v;
    dh[int64_t(b) * n + a] = v;
  }
}

// SumAClk(s,k) = sum_{i in s} C(i,k)^2 (cho_lk_scf.F90:235-253).
__global__ void k_lk_sumac(LkDens D, int nshell, const int32_t *__restrict__ koff,
                           const int32_t *__restrict__ nbsh, double *__restrict__ sumac) {
  int d, k;
  dk_of(D, blockIdx.x, &d, &k);
  const double *c = D.c[d];
  const int no = D.nocc[d];
  for (int s = threadIdx.x; s < nshell; s += blockDim.x) {
    double t = 0.0;
    for (int i = 0; i < nbsh[s]; ++i) {
      const double v = c[k + int64_t(no) * (koff[s] + i)];
      t = t + v * v;
    }
    sumac[int64_t(blockIdx.x) * nshell + s] = t;
  }
}

// Per orbital (one workgroup each): when screening, Y = DH |C_k| (DGEMV 'N',
// :584-590), MLk per shell, the listed shells; then, every batch, Keep and the
// number of half-transformed shell pairs (:659-702).
__global__ __launch_bounds__(kLkThreads) void k_lk_lists(
    LkDens D, int n, int nshell, int doscreen, int cas_cut, const int32_t *__restrict__ koff,
    const int32_t *__restrict__ nbsh, const double *__restrict__ dh, double *__restrict__ ylk,
    double *__restrict__ mlk, uint8_t *__restrict__ listed, const double *__restrict__ sumac,
    const double *__restrict__ sv, const uint8_t *__restrict__ present, uint8_t *__restrict__ keep,
    unsigned long long *__restrict__ stats) {
  __shared__ double red[kLkThreads / 64];
  __shared__ unsigned long long cnt[2];
  const int g = blockIdx.x;
  int d, k;
  dk_of(D, g, &d, &k);
  const int64_t so = int64_t(g) * nshell;
  if (threadIdx.x == 0) cnt[0] = cnt[1] = 0;
  if (doscreen) {
    const double *c = D.c[d];
    const int no = D.nocc[d];
    double *y = ylk + int64_t(g) * n;
    for (int l = threadIdx.x; l < n; l += blockDim.x) {
      double s = 0.0;
      for (int m = 0; m < n; ++m) s = s + dh[int64_t(m) * n + l] * fabs(c[k + int64_t(no) * m]);
      y[l] = s;
    }
    __syncthreads();
    double mx = 0.0;
    for (int s = threadIdx.x; s < nshell; s += blockDim.x) {
      double ym = 0.0;
      for (int i = 0; i < nbsh[s]; ++i) ym = fmax(ym, y[koff[s] + i]);
      mlk[so + s] = ym;
      mx = fmax(mx, ym);
    }
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    mx = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    unsigned long long nl = 0;
    for (int s = threadIdx.x; s < nshell; s += blockDim.x) {
      const double m = mlk[so + s];
      const bool ok = cas_cut ? (m >= D.xtau[d]) : (m * mx >= D.tau[d]);
      listed[so + s] = ok ? 1 : 0;
      nl += ok;
    }
    if (nl) atomicAdd(&cnt[0], nl);
  }
  __syncthreads();
  unsigned long long nh = 0;
  const double *sa_c = sumac + so;
  for (int sa = threadIdx.x; sa < nshell; sa += blockDim.x) {
    int count = 0;
    if (listed[so + sa])
      for (int sb = 0; sb < nshell; ++sb) {
        const int64_t p = shp_idx(sa, sb);
        if (!present[p] || nbsh[sa] * nbsh[sb] <= 0) continue;
        if (sqrt(fabs(sa_c[sb] * sv[p])) >= D.thrv[d]) ++count;
      }
    keep[so + sa] = count > 0 ? 1 : 0;
    nh += count;
  }
  if (nh) atomicAdd(&cnt[1], nh);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (cnt[0]) atomicAdd(&stats[0], cnt[0]);
    if (cnt[1]) atomicAdd(&stats[1], cnt[1]);
  }
}

// Faa(a) = max_{i in a} |sum_J X(i,J)^2| for listed, kept shells (:713-743),
// then the number of K blocks the pair cuts keep (:749-830).
__global__ __launch_bounds__(kLkThreads) void k_lk_faa(
    LkDens D, int n_pad, int jnum, int nshell, const int32_t *__restrict__ koff,
    const int32_t *__restrict__ nbsh, const double *__restrict__ X, const uint8_t *__restrict__ listed,
    const uint8_t *__restrict__ keep, const double *__restrict__ mlk, double *__restrict__ faa,
    double maxred, unsigned long long *__restrict__ stats) {
  __shared__ unsigned long long cnt;
  const int g = blockIdx.x;
  int d, k;
  dk_of(D, g, &d, &k);
  const int64_t so = int64_t(g) * nshell;
  if (threadIdx.x == 0) cnt = 0;
  const double *x = X + int64_t(g) * jnum * n_pad;
  for (int s = threadIdx.x; s < nshell; s += blockDim.x) {
    double t = 0.0;
    if (listed[so + s] && keep[so + s])
      for (int i = 0; i < nbsh[s]; ++i) {
        double f = 0.0;
        for (int j = 0; j < jnum; ++j) {
          const double v = x[int64_t(j) * n_pad + koff[s] + i];
          f = f + v * v;
        }
        t = fmax(t, fabs(f));
      }
    faa[so + s] = t;
  }
  __syncthreads();
  const int64_t nshp = int64_t(nshell) * (nshell + 1) / 2;
  unsigned long long c = 0;
  const double tau = D.tau[d], tk = D.tau[d] / maxred;
  for (int64_t p = threadIdx.x; p < nshp; p += blockDim.x) {
    int sa = int((sqrt(8.0 * double(p) + 1.0) - 1.0) * 0.5);
    while (int64_t(sa) * (sa + 1) / 2 > p) --sa;
    while (int64_t(sa + 1) * (sa + 2) / 2 <= p) ++sa;
    const int sb = int(p - int64_t(sa) * (sa + 1) / 2);
    if (!listed[so + sa] || !listed[so + sb] || !keep[so + sa] || !keep[so + sb]) continue;
    if (mlk[so + sa] * mlk[so + sb] < tau) continue;
    if (!(sqrt(fabs(faa[so + sa] * faa[so + sb])) >= tk)) continue;
    ++c;
  }
  if (c) atomicAdd(&cnt, c);
  __syncthreads();
  if (threadIdx.x == 0 && cnt) atomicAdd(&stats[2], cnt);
}

#pragma clang fp contract(fast)

// Masked half-transform of one vector: X(g, J, a) = sum_b m(g, sa, sb) L_J(a,b)
// C(b, k) for the listed shells a of orbital g = (d, k), zero elsewhere.
// Workgroup: 64 AOs a x 16 orbitals of one density, one vector J.
__global__ __launch_bounds__(kLkThreads) void k_lk_ht(
    LkDens D, const double *__restrict__ L, int64_t ldl, int64_t j0, int jnum, int n, int n_pad,
    int nshell, const int32_t *__restrict__ shell, const double *__restrict__ sumac,
    const double *__restrict__ sv, const uint8_t *__restrict__ present,
    const uint8_t *__restrict__ listed, double *__restrict__ X) {
  __shared__ double tl[kT][kT + 1];
  __shared__ double tc[kT][kKc];
  __shared__ int tsb[kT];
  const int tid = threadIdx.x, al = tid & 63, kq = tid >> 6;
  const int a0 = blockIdx.x * kT, jr = blockIdx.y;
  int d = 0, ch = blockIdx.z;
  if (D.nd > 1 && ch >= D.nkc[0]) {
    d = 1;
    ch -= D.nkc[0];
  }
  const int kbase = ch * kKc, no = D.nocc[d];
  const double *c = D.c[d];
  const double *row = L + (j0 + jr) * ldl;
  const int a = a0 + al;
  const int sa = a < n ? shell[a] : -1;
  bool lst[4];
  int kk[4];
  bool mine = false;
  for (int i = 0; i < 4; ++i) {
    kk[i] = kbase + kq + 4 * i;
    const int g = D.k0[d] + kk[i];
    lst[i] = sa >= 0 && kk[i] < no && listed[int64_t(g) * nshell + sa];
    mine |= lst[i];
  }
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (__syncthreads_or(mine)) {
    for (int b0 = 0; b0 < n; b0 += kT) {
      // stage L_J(a0.., b0..): packed rows are contiguous in the smaller index
      if (a0 > b0) {
        for (int e = tid; e < kT * kT; e += kLkThreads) {
          const int r = e >> 6, cc = e & 63;
          const int64_t aa = a0 + r, bb = b0 + cc;
          tl[r][cc] = (aa < n) ? row[aa * (aa + 1) / 2 + bb] : 0.0;
        }
      } else if (a0 < b0) {
        for (int e = tid; e < kT * kT; e += kLkThreads) {
          const int r = e & 63, cc = e >> 6;
          const int64_t aa = a0 + r, bb = b0 + cc;
          tl[r][cc] = (bb < n) ? row[bb * (bb + 1) / 2 + aa] : 0.0;
        }
      } else {
        for (int e = tid; e < kT * kT; e += kLkThreads) {
          const int r = e >> 6, cc = e & 63;
          if (cc > r) continue;
          const int64_t aa = a0 + r, bb = b0 + cc;
          const double v = (aa < n) ? row[aa * (aa + 1) / 2 + bb] : 0.0;
          tl[r][cc] = v;
          tl[cc][r] = v;
        }
      }
      for (int e = tid; e < kT * kKc; e += kLkThreads) {
        const int r = e / kKc, q = e % kKc;
        const int64_t bb = b0 + r;
        const int k = kbase + q;
        tc[r][q] = (bb < n && k < no) ? c[k + int64_t(no) * bb] : 0.0;
      }
      if (tid < kT) tsb[tid] = (b0 + tid < n) ? shell[b0 + tid] : -1;
      __syncthreads();
      if (mine) {
        int prev = -2;
        bool m[4] = {false, false, false, false};
        const int bend = min(kT, n - b0);
        for (int cc = 0; cc < bend; ++cc) {
          const int sb = tsb[cc];
          if (sb != prev) {
            prev = sb;
            const int64_t p = shp_idx(sa, sb);
            const bool pres = present[p] != 0;
            const double svv = sv[p];
            for (int i = 0; i < 4; ++i) {
              const int g = D.k0[d] + kk[i];
              m[i] = lst[i] && pres && sqrt(fabs(sumac[int64_t(g) * nshell + sb] * svv)) >= D.thrv[d];
            }
          }
          const double l = tl[al][cc];
          for (int i = 0; i < 4; ++i)
            if (m[i]) acc[i] += l * tc[cc][kq + 4 * i];
        }
      }
      __syncthreads();
    }
  }
  if (a < n_pad)
    for (int i = 0; i < 4; ++i)
      if (kk[i] < no) X[(int64_t(D.k0[d] + kk[i]) * jnum + jr) * n_pad + a] = acc[i];
}

// K tile (A, B), A >= B, of the lower triangle: kacc[d](a,b) += sum over the
// orbitals k of density d whose pair cuts keep (sa, sb) of sum_J X X.
__global__ __launch_bounds__(kLkThreads) void k_lk_k(
    LkDens D, const double *__restrict__ X, int jnum, int n, int n_pad, int nshell,
    const int32_t *__restrict__ shell, const uint8_t *__restrict__ listed,
    const uint8_t *__restrict__ keep, const double *__restrict__ mlk, const double *__restrict__ faa,
    double maxred, double *__restrict__ kacc, int64_t ldk) {
  __shared__ double xa[kLkBatch][kT];
  __shared__ double xb[kLkBatch][kT];
  // tile (ta, tb), ta >= tb, from the linear block index
  const int64_t t = blockIdx.x;
  int ta = int((sqrt(8.0 * double(t) + 1.0) - 1.0) * 0.5);
  while (int64_t(ta) * (ta + 1) / 2 > t) --ta;
  while (int64_t(ta + 1) * (ta + 2) / 2 <= t) ++ta;
  const int tb = int(t - int64_t(ta) * (ta + 1) / 2);
  const int tid = threadIdx.x, ti = tid & 15, tj = tid >> 4;
  int ar[4], bc[4], sa[4], sb[4];
  for (int i = 0; i < 4; ++i) {
    ar[i] = ta * kT + ti * 4 + i;
    bc[i] = tb * kT + tj * 4 + i;
    sa[i] = ar[i] < n ? shell[ar[i]] : -1;
    sb[i] = bc[i] < n ? shell[bc[i]] : -1;
  }
  for (int d = 0; d < D.nd; ++d) {
    double acc[4][4] = {};
    const double tau = D.tau[d], tk = D.tau[d] / maxred;
    for (int k = 0; k < D.nocc[d]; ++k) {
      const int g = D.k0[d] + k;
      const int64_t so = int64_t(g) * nshell;
      bool m[4][4];
      bool any = false;
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
          bool ok = sa[i] >= 0 && sb[j] >= 0 && ar[i] >= bc[j];
          if (ok) {
            const int s1 = sa[i], s2 = sb[j];
            ok = listed[so + s1] && listed[so + s2] && keep[so + s1] && keep[so + s2] &&
                 !(mlk[so + s1] * mlk[so + s2] < tau) &&
                 sqrt(fabs(faa[so + s1] * faa[so + s2])) >= tk;
          }
          m[i][j] = ok;
          any |= ok;
        }
      if (!__syncthreads_or(any)) continue;
      const double *xg = X + int64_t(g) * jnum * n_pad;
      for (int e = tid; e < jnum * kT; e += kLkThreads) {
        const int j = e >> 6, cc = e & 63;
        xa[j][cc] = (ta * kT + cc < n_pad) ? xg[int64_t(j) * n_pad + ta * kT + cc] : 0.0;
        xb[j][cc] = (tb * kT + cc < n_pad) ? xg[int64_t(j) * n_pad + tb * kT + cc] : 0.0;
      }
      __syncthreads();
      if (any) {
        double s[4][4] = {};
        for (int j = 0; j < jnum; ++j) {
          double va[4], vb[4];
          for (int i = 0; i < 4; ++i) {
            va[i] = xa[j][ti * 4 + i];
            vb[i] = xb[j][tj * 4 + i];
          }
          for (int i = 0; i < 4; ++i)
            for (int q = 0; q < 4; ++q) s[i][q] += va[i] * vb[q];
        }
        for (int i = 0; i < 4; ++i)
          for (int q = 0; q < 4; ++q)
            if (m[i][q]) acc[i][q] += s[i][q];
      }
      __syncthreads();
    }
    double *kd = kacc + int64_t(d) * ldk;
    for (int i = 0; i < 4; ++i)
      for (int q = 0; q < 4; ++q)
        if (ar[i] < n && bc[q] < n && ar[i] >= bc[q])
          kd[int64_t(ar[i]) * (ar[i] + 1) / 2 + bc[q]] += acc[i][q];
  }
}

// Shell-pair occupancy of a reduced set (nnBstRSh > 0 of the reference).
int lk_present(Context *c, int rs, const uint8_t **out) {
  LKState &K = c->lk;
  const int ns = K.nshell;
  const int64_t nshp = int64_t(ns) * (ns + 1) / 2;
  std::vector<int32_t> shell(size_t(c->n));
  for (int s = 0, a = 0; s < ns; ++s)
    for (int i = 0; i < K.nbassh[size_t(s)]; ++i) shell[size_t(a++)] = s;
  auto lt_pair = [](int64_t idx, int64_t *a, int64_t *b) {
    int64_t r = int64_t((std::sqrt(8.0 * double(idx) + 1.0) - 1.0) * 0.5);
    while (r * (r + 1) / 2 > idx) --r;
    while ((r + 1) * (r + 2) / 2 <= idx) ++r;
    *a = r;
    *b = idx - r * (r + 1) / 2;
  };
  DevBuf *buf;
  int64_t *ver;
  std::vector<uint8_t> h(size_t(nshp), 0);
  if (rs < 0) {  // every pair: the store written in place
    buf = &K.present_all;
    ver = &K.present_all_version;
    if (*ver != K.version) std::fill(h.begin(), h.end(), 1);
  } else {
    if (K.present.size() < c->rsets.size()) {
      K.present.resize(c->rsets.size());
      K.present_version.resize(c->rsets.size(), -1);
    }
    buf = &K.present[size_t(rs)];
    ver = &K.present_version[size_t(rs)];
    if (*ver != K.version) {
      const ReducedSet &R = c->rsets[size_t(rs)];
      for (int64_t r = 0; r < R.nrs; ++r) {
        int64_t a, b;
        lt_pair(R.identity ? r : int64_t(R.h_map[size_t(r)]), &a, &b);
        const int s1 = shell[size_t(a)], s2 = shell[size_t(b)];
        h[size_t(s1 >= s2 ? int64_t(s1) * (s1 + 1) / 2 + s2 : int64_t(s2) * (s2 + 1) / 2 + s1)] = 1;
      }
    }
  }
  if (*ver != K.version) {
    int rc;
    if ((rc = ensure(c, *buf, size_t(nshp)))) return rc;
    if ((rc = hip_check(c, hipMemcpy(buf->ptr, h.data(), size_t(nshp), hipMemcpyHostToDevice),
                        "LK shell-pair occupancy")))
      return rc;
    *ver = K.version;
  }
  *out = static_cast<const uint8_t *>(buf->ptr);
  return MJK_OK;
}

}  // namespace

int lk_exchange(Context *c, int32_t nd, const int32_t *nocc, const double *const *corb,
                const double *tau, const double *thrv, bool cas_cut, double *klt, int64_t ldk) {
  LKState &K = c->lk;
  if (K.nshell < 1) return set_error(c, MJK_ERR_STATE, "LK screening: no shells (mjk_set_shells)");
  if (!K.diag_set)
    return set_error(c, MJK_ERR_STATE, "LK screening: no integral diagonal (mjk_set_lk_diag)");
  if (nd < 1 || nd > 2) return set_error(c, MJK_ERR_ARG, "LK screening: one or two densities");
  if (c->nloaded != c->nchol)
    return set_error(c, MJK_ERR_SHORT_READ, "LK screening: not all Cholesky vectors were uploaded");
  const int n = c->n, n_pad = c->n_pad, ns = K.nshell;
  const int64_t np = c->np, nshp = int64_t(ns) * (ns + 1) / 2;
  // vectors grouped by reduced set, in increasing order (InfVec(J,2) of the reference)
  struct Run { int rs; int64_t j0, jn; };
  std::vector<Run> runs;
  for (int64_t j = 0; j < c->nchol; ++j) {
    const int rs = c->vset[size_t(j)] > 0 ? c->vset[size_t(j)] - 1 : -1;
    if (runs.empty() || runs.back().rs != rs) runs.push_back({rs, j, 0});
    runs.back().jn += 1;
  }
  for (size_t i = 1; i < runs.size(); ++i)
    if (runs[i].rs < 0 || runs[i - 1].rs < 0 || runs[i].rs <= runs[i - 1].rs)
      return set_error(c, MJK_ERR_STATE,
                       "LK screening: vectors must come reduced set by reduced set, in order");
  LkDens D{};
  D.nd = nd;
  int ktot = 0;
  for (int d = 0; d < nd; ++d) {
    D.nocc[d] = std::max(0, nocc[d]);
    D.k0[d] = ktot;
    D.nkc[d] = int(ceil_div(D.nocc[d], kKc));
    D.c[d] = corb[d];
    D.tau[d] = tau[d];
    D.thrv[d] = thrv[d];
    D.xtau[d] = std::sqrt(tau[d]);  // xtau = sqrt(tau), cho_lk_casscf.F90:193-194
    ktot += D.nocc[d];
  }
  D.ktot = ktot;
  int rc;
  for (int d = 0; d < nd; ++d)
    if ((rc = fill_zero(c, klt + d * ldk, sizeof(double) * size_t(np)))) return rc;
  K.last_stats[0] = K.last_stats[1] = K.last_stats[2] = 0;
  K.last_batches = 0;
  if (ktot == 0 || runs.empty()) return MJK_OK;
  const int64_t kt = std::max(ktot, 1);
  const size_t flag_bytes = size_t(round_up(kt * ns, 4));  // fill_zero works in 4-byte words
  if ((rc = ensure(c, K.diag, sizeof(double) * size_t(np))) ||
      (rc = ensure(c, K.dh, sizeof(double) * size_t(n) * n)) ||
      (rc = ensure(c, K.ylk, sizeof(double) * size_t(kt) * n)) ||
      (rc = ensure(c, K.svshp, sizeof(double) * size_t(nshp))) ||
      (rc = ensure(c, K.sumac, sizeof(double) * size_t(kt) * ns)) ||
      (rc = ensure(c, K.mlk, sizeof(double) * size_t(kt) * ns)) ||
      (rc = ensure(c, K.faa, sizeof(double) * size_t(kt) * ns)) ||
      (rc = ensure(c, K.listed, flag_bytes)) || (rc = ensure(c, K.keep, flag_bytes)) ||
      (rc = ensure(c, K.x, sizeof(double) * size_t(kt) * kLkBatch * n_pad)) ||
      (rc = ensure(c, K.stats, 3 * sizeof(unsigned long long))))
    return rc;
  double *diag = static_cast<double *>(K.diag.ptr);
  if ((rc = hip_check(c, hipMemcpyAsync(diag, K.diag0.ptr, sizeof(double) * size_t(np),
                                        hipMemcpyDeviceToDevice, c->stream), "LK diagonal")) ||
      (rc = fill_zero(c, K.stats.ptr, 3 * sizeof(unsigned long long))) ||
      (rc = fill_zero(c, K.faa.ptr, sizeof(double) * size_t(kt) * ns)) ||
      (rc = fill_zero(c, K.listed.ptr, flag_bytes)) ||
      (rc = fill_zero(c, K.mlk.ptr, sizeof(double) * size_t(kt) * ns)))
    return rc;
  const int32_t *shell = static_cast<const int32_t *>(K.shell.ptr);
  const int32_t *koff = static_cast<const int32_t *>(K.koff.ptr);
  const int32_t *nbsh = static_cast<const int32_t *>(K.nbsh.ptr);
  double *sv = static_cast<double *>(K.svshp.ptr), *sumac = static_cast<double *>(K.sumac.ptr);
  double *mlk = static_cast<double *>(K.mlk.ptr), *faa = static_cast<double *>(K.faa.ptr);
  uint8_t *listed = static_cast<uint8_t *>(K.listed.ptr), *keep = static_cast<uint8_t *>(K.keep.ptr);
  double *X = static_cast<double *>(K.x.ptr);
  auto *stats = static_cast<unsigned long long *>(K.stats.ptr);
  const double maxred = double(K.maxred > 0 ? K.maxred : int(std::max<size_t>(c->rsets.size(), 1)));
  hipLaunchKernelGGL(k_lk_sumac, dim3(unsigned(ktot)), dim3(kLkThreads), 0, c->stream, D, ns, koff,
                     nbsh, sumac);
  const unsigned gnp = unsigned(std::min<int64_t>(ceil_div(np, kLkThreads), 4096));
  const int ntile = int(ceil_div(n, kT));
  const unsigned nkt = unsigned(int64_t(ntile) * (ntile + 1) / 2);
  // reduced sets JRED1..JRED2 of the vectors; the screening cadence counts
  // every one of them, with or without vectors (:336-337,891-898)
  const int jred1 = runs.front().rs, jred2 = runs.back().rs;
  int kscreen = 1;
  bool doscreen = true;
  size_t ri = 0;
  for (int jred = jred1; jred <= jred2; ++jred) {
    if (ri < runs.size() && runs[ri].rs == jred) {
      const Run &R = runs[ri++];
      const uint8_t *present = nullptr;
      if ((rc = lk_present(c, R.rs, &present))) return rc;
      const int64_t nvec = std::min<int64_t>(kLkBatch, R.jn);  // :374
      for (int64_t b = 0; b < R.jn; b += nvec) {
        const int64_t j0 = R.j0 + b;
        const int jnum = int(std::min<int64_t>(nvec, R.jn - b));
        if (K.estimate)
          hipLaunchKernelGGL(k_lk_est_diag, dim3(gnp), dim3(kLkThreads), 0, c->stream, c->L, c->ldl,
                             j0, jnum, np, diag);
        hipLaunchKernelGGL(k_lk_svshp, dim3(unsigned(ceil_div(nshp, kLkThreads / 64))),
                           dim3(kLkThreads), 0, c->stream, c->L, c->ldl, j0, jnum, ns, koff, nbsh, sv);
        if (doscreen)
          hipLaunchKernelGGL(k_lk_diah, dim3(unsigned(n)), dim3(kLkThreads), 0, c->stream, diag, n,
                             static_cast<double *>(K.dh.ptr));
        hipLaunchKernelGGL(k_lk_lists, dim3(unsigned(ktot)), dim3(kLkThreads), 0, c->stream, D, n, ns,
                           int(doscreen), int(cas_cut), koff, nbsh,
                           static_cast<const double *>(K.dh.ptr), static_cast<double *>(K.ylk.ptr),
                           mlk, listed, sumac, sv, present, keep, stats);
        hipLaunchKernelGGL(k_lk_ht, dim3(unsigned(ntile), unsigned(jnum), unsigned(D.nkc[0] + D.nkc[1])),
                           dim3(kLkThreads), 0, c->stream, D, c->L, c->ldl, j0, jnum, n, n_pad, ns,
                           shell, sumac, sv, present, listed, X);
        hipLaunchKernelGGL(k_lk_faa, dim3(unsigned(ktot)), dim3(kLkThreads), 0, c->stream, D, n_pad,
                           jnum, ns, koff, nbsh, X, listed, keep, mlk, faa, maxred, stats);
        hipLaunchKernelGGL(k_lk_k, dim3(nkt), dim3(kLkThreads), 0, c->stream, D, X, jnum, n, n_pad, ns,
                           shell, listed, keep, mlk, faa, maxred, klt, ldk);
        if (K.update && c->lk_nscreen > 0)
          hipLaunchKernelGGL(k_lk_update_diag, dim3(gnp), dim3(kLkThreads), 0, c->stream, c->L,
                             c->ldl, j0, jnum, np, diag);
        if ((rc = hip_check(c, hipGetLastError(), "LK batch launch"))) return rc;
        doscreen = false;  // :831
        ++K.last_batches;
      }
    }
    doscreen = kscreen == c->lk_nscreen;  // :891-898
    if (!doscreen) ++kscreen;
    else kscreen = 1;
  }
  unsigned long long h[3];
  if ((rc = hip_check(c, hipMemcpyAsync(h, stats, sizeof(h), hipMemcpyDeviceToHost, c->stream),
                      "LK counters")) ||
      (rc = hip_check(c, hipStreamSynchronize(c->stream), "LK sync")))
    return rc;
  for (int i = 0; i < 3; ++i) K.last_stats[i] = (long long)h[i];
  return MJK_OK;
}

}  // namespace mjk
