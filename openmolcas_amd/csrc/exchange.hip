// Exchange (K) build: half-transform X^J = C^T L^J and K = sum_J X^J^T X^J.
//
// Reference: Cho_X_getVtra -> CHO_VTRA iSwap=2 (src/cholesky_util/cho_vtra.F90:
// 318-363), a scatter-AXPY over reduced-set pairs, followed by DGEMM_Tri
// (src/linalg_util/dgemm_tri.F90:122-147), a lower-triangle SYRK over the merged
// (k,J) index executed as per-column DGEMV (src/fock_util/cho_fscf.F90:276).
//
// MI355X design:
//  * All orbital sets of one call are concatenated into one "orbital bank"
//    Cb[b][k] (AO row b, orbital column k; every set starts at a multiple of 4
//    columns, padded with zeros), so ONE pass over L serves every density.
//  * Half-transform: a workgroup owns (vector J, 64-wide AO tile a0) and all
//    bank rows k.  It walks b in chunks of BK, unpacks the packed-triangle
//    block L^J[b0:b0+BK, a0:a0+64] into a square LDS tile (mirroring across
//    the diagonal; the reference's 1/2 on diagonal pairs added twice is the
//    same as taking the diagonal once), stages Cb[b0:b0+BK, :] in LDS and
//    runs v_mfma_f64_16x16x4f64 with wave w owning AO columns 16w..16w+15.
//    Output X[J][k][a] (a fastest).  Workgroups of one vector are placed on
//    one XCD so that the two reads of every packed element (as row and as
//    transposed column) share that XCD's L2.
//  * SYRK: a workgroup owns one TMxTM tile (A >= B) of K and a slice of the
//    vectors; it contracts over (J, k in the density's bank columns) with
//    v_mfma_f64_16x16x4f64 and writes an FP64 partial tile.  A final
//    deterministic reduction sums the slices into LT storage.
#include "mjk_internal.h"

namespace mjk {
namespace {

constexpr int kThreads = 256;

__device__ inline v4d mfma_f64(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ inline int64_t tri0(int64_t a) { return a * (a + 1) / 2; }

// Maps blockIdx.x so that consecutive logical blocks land on the same XCD
// (hardware dispatches block i to XCD i % 8).  Returns -1 for padding blocks.
__device__ inline int64_t xcd_logical(int64_t nblocks) {
  const int64_t g = gridDim.x;  // multiple of 8
  const int64_t b = blockIdx.x;
  const int64_t l = (b & 7) * (g >> 3) + (b >> 3);
  return l < nblocks ? l : -1;
}

// ---- orbital bank --------------------------------------------------------
// Cb[a*ldc + off + k] = corb[k + a*nocc] for a < n, k < nocc.
__global__ void k_pack_orb(const double *__restrict__ corb, int nocc, int n, int ldc, int off,
                           double *__restrict__ cb) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= int64_t(nocc) * n) return;
  const int64_t a = i / nocc, k = i % nocc;
  cb[a * ldc + off + k] = corb[i];
}

// ---- half-transform ------------------------------------------------------
// grid.x = round_up(jc * ntile_a, 8), block = 256.
// X[(jr*ldc + k)*n_pad + a] = sum_b Cb[b*ldc + k] * L^{j0+jr}_{ab}.
template <int OT, int BK>
__global__ __launch_bounds__(kThreads, 2) void k_ht_f64(
    const double *__restrict__ L, int64_t ldl, int n, int n_pad, const double *__restrict__ Cb,
    int64_t j0, int ntile_a, int64_t nblocks, double *__restrict__ X) {
  constexpr int LDC = 16 * OT;
  __shared__ double sC[BK][LDC];
  __shared__ double sL[BK][kTileA + 1];
  const int64_t lb = xcd_logical(nblocks);
  if (lb < 0) return;
  const int64_t jr = lb / ntile_a;
  const int a0 = int(lb % ntile_a) * kTileA;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const double *__restrict__ Lj = L + (j0 + jr) * ldl;

  v4d acc[OT];
#pragma unroll
  for (int t = 0; t < OT; ++t) acc[t] = v4d{0.0, 0.0, 0.0, 0.0};

  if (a0 < n) {
    for (int b0 = 0; b0 < n; b0 += BK) {
      // stage bank rows b0..b0+BK-1 (contiguous, zero padded beyond n)
      const double *__restrict__ cs = Cb + int64_t(b0) * LDC;
#pragma unroll
      for (int i = tid; i < BK * LDC; i += kThreads) sC[i / LDC][i % LDC] = cs[i];
      // unpack L^J[b0:b0+BK, a0:a0+64] into sL[b][a]
      if (b0 >= a0 + kTileA) {  // every b > a: row b holds a contiguous
#pragma unroll
        for (int i = tid; i < BK * kTileA; i += kThreads) {
          const int bb = i / kTileA, aa = i % kTileA, b = b0 + bb, a = a0 + aa;
          sL[bb][aa] = (b < n && a < n) ? Lj[tri0(b) + a] : 0.0;
        }
      } else if (b0 + BK <= a0) {  // every b < a: row a holds b contiguous
#pragma unroll
        for (int i = tid; i < BK * kTileA; i += kThreads) {
          const int bb = i % BK, aa = i / BK, b = b0 + bb, a = a0 + aa;
          sL[bb][aa] = (a < n) ? Lj[tri0(a) + b] : 0.0;
        }
      } else {  // block straddles the diagonal
#pragma unroll
        for (int i = tid; i < BK * kTileA; i += kThreads) {
          const int bb = i / kTileA, aa = i % kTileA, b = b0 + bb, a = a0 + aa;
          double v = 0.0;
          if (b < n && a < n) v = (b >= a) ? Lj[tri0(b) + a] : Lj[tri0(a) + b];
          sL[bb][aa] = v;
        }
      }
      __syncthreads();
#pragma unroll
      for (int ks = 0; ks < BK / 4; ++ks) {
        const int kr = ks * 4 + (lane >> 4);
        const double bf = sL[kr][16 * w + (lane & 15)];
#pragma unroll
        for (int t = 0; t < OT; ++t) acc[t] = mfma_f64(sC[kr][16 * t + (lane & 15)], bf, acc[t]);
      }
      __syncthreads();
    }
  }
  // D[i][j] of v_mfma_f64_16x16x4f64 on gfx950: register r of lane l holds
  // row i = 4*r + l/16, column j = l%16.
  const int a = a0 + 16 * w + (lane & 15);
  double *__restrict__ xj = X + jr * LDC * int64_t(n_pad);
#pragma unroll
  for (int t = 0; t < OT; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) xj[int64_t(16 * t + 4 * r + (lane >> 4)) * n_pad + a] = acc[t][r];
}

// ---- SYRK ------------------------------------------------------------------
// grid.x = ntiles * nsplit.  Tile t = A(A+1)/2 + B (A >= B) of TMxTM, split s
// covers chunk vectors [s*jper, min(jc,(s+1)*jper)).  The contraction runs
// over f = (J, k), k in [off, off+opad) of the bank columns.
// kpart[(s*ntiles + t)*TM*TM + al*TM + bl] (+)= sum_f X[f][A*TM+al] X[f][B*TM+bl]
template <int TM>
__global__ __launch_bounds__(kThreads, 2) void k_syrk_f64(
    const double *__restrict__ X, int n_pad, int ldc, int off, int opad, int64_t jc,
    int64_t jper, int ntiles, double *__restrict__ kpart, int beta) {
  constexpr int BKF = 32;
  constexpr int WT = TM / 2;
  constexpr int MT = WT / 16;
  __shared__ double sA[BKF][TM];
  __shared__ double sB[BKF][TM];
  __shared__ int64_t srow[BKF];
  const int tile = blockIdx.x % ntiles;
  const int s = blockIdx.x / ntiles;
  int A = int((sqrt(8.0 * tile + 1.0) - 1.0) * 0.5);
  while (A * (A + 1) / 2 > tile) --A;
  while ((A + 1) * (A + 2) / 2 <= tile) ++A;
  const int B = tile - A * (A + 1) / 2;
  const int64_t ja = int64_t(s) * jper;
  const int64_t jz = min(jc, ja + jper);
  const int64_t nf = (jz > ja ? jz - ja : 0) * opad;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;

  v4d acc[MT][MT];
#pragma unroll
  for (int i = 0; i < MT; ++i)
#pragma unroll
    for (int j = 0; j < MT; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

  for (int64_t f0 = 0; f0 < nf; f0 += BKF) {
    // X row offset of each contraction index f = (J, k) of this chunk
    if (tid < BKF) {
      const int64_t f = f0 + tid;
      int64_t ro = -1;
      if (f < nf) {
        const int fi = int(f);  // nf < 2^31 (checked on the host)
        ro = ((ja + fi / opad) * ldc + off + fi % opad) * int64_t(n_pad);
      }
      srow[tid] = ro;
    }
    __syncthreads();
#pragma unroll
    for (int i = tid; i < BKF * TM; i += kThreads) {
      const int ff = i / TM, cc = i % TM;
      const int64_t ro = srow[ff];
      double va = 0.0, vb = 0.0;
      if (ro >= 0) {
        va = X[ro + A * TM + cc];
        vb = X[ro + B * TM + cc];
      }
      sA[ff][cc] = va;
      sB[ff][cc] = vb;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < BKF / 4; ++ks) {
      const int kr = ks * 4 + (lane >> 4);
      double af[MT], bf[MT];
#pragma unroll
      for (int i = 0; i < MT; ++i) af[i] = sA[kr][wm * WT + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < MT; ++j) bf[j] = sB[kr][wn * WT + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < MT; ++j) acc[i][j] = mfma_f64(af[i], bf[j], acc[i][j]);
    }
    __syncthreads();
  }
  double *__restrict__ kp = kpart + (int64_t(s) * ntiles + tile) * TM * TM;
#pragma unroll
  for (int i = 0; i < MT; ++i)
#pragma unroll
    for (int j = 0; j < MT; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int al = wm * WT + i * 16 + 4 * r + (lane >> 4);
        const int bl = wn * WT + j * 16 + (lane & 15);
        double *dst = kp + al * TM + bl;
        *dst = beta ? *dst + acc[i][j][r] : acc[i][j][r];
      }
}

// klt[a*(a+1)/2 + b] = sum_s kpart[(s*ntiles + tile(a,b))*TM*TM + (a%TM)*TM + b%TM]
// grid.x = n (row a), block 256.
__global__ void k_reduce_k(const double *__restrict__ kpart, int n, int tm, int ntiles,
                           int nsplit, double *__restrict__ klt) {
  const int a = blockIdx.x;
  const int A = a / tm, al = a % tm;
  double *__restrict__ out = klt + tri0(a);
  for (int b = threadIdx.x; b <= a; b += blockDim.x) {
    const int B = b / tm, bl = b % tm;
    const int64_t t = int64_t(A) * (A + 1) / 2 + B;
    const double *__restrict__ src = kpart + t * tm * tm + al * tm + bl;
    const int64_t stride = int64_t(ntiles) * tm * tm;
    double s = 0.0;
    for (int k = 0; k < nsplit; ++k) s += src[k * stride];
    out[b] = s;
  }
}

template <int OT>
void launch_ht(Context *c, const double *cb, int64_t j0, int64_t jc, double *X) {
  const int ntile_a = c->n_pad / kTileA;
  const int64_t nblocks = jc * ntile_a;
  const int64_t grid = round_up(nblocks, 8);
  constexpr int BK = OT <= 10 ? 32 : 16;
  hipLaunchKernelGGL((k_ht_f64<OT, BK>), dim3(unsigned(grid)), dim3(kThreads), 0, c->stream, c->L,
                     c->ldl, c->n, c->n_pad, cb, j0, ntile_a, nblocks, X);
}


}  // namespace

void ht_launch(Context *c, int ot, const double *cb, int64_t j0, int64_t jn, double *X) {
  switch (ot) {
    case 1: launch_ht<1>(c, cb, j0, jn, X); break;
    case 2: launch_ht<2>(c, cb, j0, jn, X); break;
    case 3: launch_ht<3>(c, cb, j0, jn, X); break;
    case 4: launch_ht<4>(c, cb, j0, jn, X); break;
    case 6: launch_ht<6>(c, cb, j0, jn, X); break;
    case 8: launch_ht<8>(c, cb, j0, jn, X); break;
    case 10: launch_ht<10>(c, cb, j0, jn, X); break;
    case 12: launch_ht<12>(c, cb, j0, jn, X); break;
    default: launch_ht<16>(c, cb, j0, jn, X); break;
  }
}

int pack_bank(Context *c, int32_t nk, const int32_t *nocc, const double *const *corb,
              std::vector<int> &off, std::vector<int> &opad, int *ot_out, int *tot_out) {
  off.assign(size_t(nk), 0);
  opad.assign(size_t(nk), 0);
  int tot = 0;
  for (int d = 0; d < nk; ++d) {
    if (nocc[d] < 0 || nocc[d] > 256)
      return set_error(c, MJK_ERR_ARG, "orbital set larger than 256 columns");
    off[size_t(d)] = tot;
    opad[size_t(d)] = int(round_up(nocc[d], 4));
    tot += opad[size_t(d)];
  }
  *tot_out = tot;
  *ot_out = 0;
  if (tot == 0) return MJK_OK;
  static const int kOT[] = {1, 2, 3, 4, 6, 8, 10, 12, 16};
  int ot = 0;
  for (int v : kOT)
    if (16 * v >= tot) {
      ot = v;
      break;
    }
  if (!ot) return set_error(c, MJK_ERR_ARG, "more than 256 orbital columns in one call");
  const int ldc = 16 * ot;
  const int n = c->n, n_pad = c->n_pad;
  int rc;
  if ((rc = ensure(c, c->ws_cbank, sizeof(double) * int64_t(n_pad) * ldc))) return rc;
  double *cb = static_cast<double *>(c->ws_cbank.ptr);
  phase_begin(c);
  if ((rc = hip_check(c, hipMemsetAsync(cb, 0, sizeof(double) * int64_t(n_pad) * ldc, c->stream),
                      "memset bank")))
    return rc;
  for (int d = 0; d < nk; ++d) {
    if (nocc[d] == 0) continue;
    const int64_t cnt = int64_t(nocc[d]) * n;
    hipLaunchKernelGGL(k_pack_orb, dim3(unsigned(ceil_div(cnt, 256))), dim3(256), 0, c->stream,
                       corb[d], nocc[d], n, ldc, off[size_t(d)], cb);
  }
  phase_end(c, MJK_PHASE_PACK);
  *ot_out = ot;
  return MJK_OK;
}

int exchange(Context *c, int32_t nk, const int32_t *nocc, const double *const *corb, double *klt,
             int64_t ldk, int32_t precision) {
  if (nk < 1) return set_error(c, MJK_ERR_ARG, "exchange: nk < 1");
  if (ldk < c->np) return set_error(c, MJK_ERR_ARG, "exchange: ldk < nbas(nbas+1)/2");
  if (precision != MJK_PREC_FP64)
    return set_error(c, MJK_ERR_ARG, "exchange: precision mode not available");
  if (c->nloaded != c->nchol)
    return set_error(c, MJK_ERR_SHORT_READ, "exchange: not all Cholesky vectors were uploaded");
  const int n = c->n, n_pad = c->n_pad;
  const int64_t M = c->nchol;
  int rc;

  // ---- orbital bank (all sets of this call, one pass over L)
  std::vector<int> off, opad;
  int ot = 0, tot = 0;
  if ((rc = pack_bank(c, nk, nocc, corb, off, opad, &ot, &tot))) return rc;
  if (tot == 0) {  // no orbitals: K = 0
    for (int d = 0; d < nk; ++d)
      if ((rc = hip_check(c, hipMemsetAsync(klt + d * ldk, 0, sizeof(double) * c->np, c->stream),
                          "memset K")))
        return rc;
    return MJK_OK;
  }
  const int ldc = 16 * ot;
  const double *cb = static_cast<const double *>(c->ws_cbank.ptr);

  // ---- chunking over vectors (X budget)
  const int64_t xvec_bytes = sizeof(double) * int64_t(ldc) * n_pad;
  int64_t jc = std::max<int64_t>(1, std::min<int64_t>(M, c->x_budget / xvec_bytes));
  const int64_t nchunk = ceil_div(M, jc);
  jc = ceil_div(M, nchunk);
  if (jc * ldc >= (int64_t(1) << 31))
    return set_error(c, MJK_ERR_ARG, "exchange: chunk contraction length exceeds 2^31");
  if ((rc = ensure(c, c->ws_x, size_t(xvec_bytes) * jc))) return rc;
  double *X = static_cast<double *>(c->ws_x.ptr);

  // ---- SYRK tiling and slices
  int tm = c->tile_syrk ? c->tile_syrk : (n_pad >= 512 && n_pad % 128 == 0 ? 128 : 64);
  if (n_pad % tm) tm = 64;
  const int ntile1 = n_pad / tm;
  const int ntiles = ntile1 * (ntile1 + 1) / 2;
  std::vector<int> nsplit(nk);
  std::vector<int64_t> kbase(nk + 1, 0);
  int active = 0;
  for (int d = 0; d < nk; ++d) active += nocc[d] > 0;
  for (int d = 0; d < nk; ++d) {
    int64_t want = ceil_div(1024, int64_t(ntiles) * std::max(active, 1));
    nsplit[d] = nocc[d] > 0 ? int(std::max<int64_t>(1, std::min<int64_t>(want, jc))) : 0;
    kbase[d + 1] = kbase[d] + int64_t(nsplit[d]) * ntiles * tm * tm;
  }
  if ((rc = ensure(c, c->ws_kpart, sizeof(double) * std::max<int64_t>(kbase[nk], 1)))) return rc;
  double *kpart = static_cast<double *>(c->ws_kpart.ptr);

  for (int64_t ch = 0; ch < nchunk; ++ch) {
    const int64_t j0 = ch * jc;
    const int64_t jn = std::min(jc, M - j0);
    phase_begin(c);
    ht_launch(c, ot, cb, j0, jn, X);
    phase_end(c, MJK_PHASE_HALF_TRANSFORM);
    phase_begin(c);
    for (int d = 0; d < nk; ++d) {
      if (!nsplit[d]) continue;
      const int64_t jper = ceil_div(jn, nsplit[d]);
      const dim3 grid(unsigned(ntiles * nsplit[d]));
      if (tm == 128)
        hipLaunchKernelGGL(k_syrk_f64<128>, grid, dim3(kThreads), 0, c->stream, X, n_pad, ldc,
                           off[d], opad[d], jn, jper, ntiles, kpart + kbase[d], ch > 0 ? 1 : 0);
      else
        hipLaunchKernelGGL(k_syrk_f64<64>, grid, dim3(kThreads), 0, c->stream, X, n_pad, ldc,
                           off[d], opad[d], jn, jper, ntiles, kpart + kbase[d], ch > 0 ? 1 : 0);
    }
    phase_end(c, MJK_PHASE_EXCHANGE);
  }
  phase_begin(c);
  for (int d = 0; d < nk; ++d) {
    if (!nsplit[d]) {
      if ((rc = hip_check(c, hipMemsetAsync(klt + d * ldk, 0, sizeof(double) * c->np, c->stream),
                          "memset K")))
        return rc;
      continue;
    }
    hipLaunchKernelGGL(k_reduce_k, dim3(unsigned(n)), dim3(256), 0, c->stream, kpart + kbase[d], n,
                       tm, ntiles, nsplit[d], klt + d * ldk);
  }
  phase_end(c, MJK_PHASE_REDUCE);
  return hip_check(c, hipGetLastError(), "exchange launch");
}

}  // namespace mjk
