// Backward of the closed-shell energy functional of the J/K build.
//
//   E(C, L) = tr(P h) + 1/2 sum_J V_J^2 - 1/4 sum_J tr(P L^J P L^J),
//   P = 2 C C^T,  V_J = tr(P L^J),  F = h + J[P] - 1/2 K[P]
//   dE/dC   = 4 F C
//   dE/dL^J = V_J P - 1/2 P L^J P = V_J P - 2 C Z^J C^T,  Z^J = C^T L^J C
// with the packed-parameter rule: off-diagonal entries of dE/dL_packed carry
// a factor 2 (L_ab and L_ba are the same parameter).
//
// The energy expression is that of src/scf/eneclc.F90:58,77,104 with the
// Fock matrix of src/fock_util/cho_fscf.F90 (DECO convention, ExFac = 1/2 on
// V = sqrt(2) C); the gradient has no counterpart in the reference, which only
// evaluates the Fock matrix.
//
// Kernels: the half-transform X^J = C^T L^J is recomputed chunk by chunk (the
// chunk size follows the context's X budget); Z = X C, W = C Z and the packed
// lower triangle of W C^T go through one batched FP64-MFMA GEMM whose epilogue
// writes the vector gradient directly in packed storage.
#include <type_traits>

#include "mjk_internal.h"

namespace mjk {
namespace {

constexpr int kThreads = 256;
constexpr int kTM = 64;  // output tile
constexpr int kBK = 16;  // contraction chunk

__device__ inline v4d mfma64(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline int64_t tri0(int64_t a) { return a * (a + 1) / 2; }

struct GemmBatch {
  const double *A;  // A[b][i][k] at A + b*sA + i*lda + k
  const float *A32; // the same for an FP32 A operand (k_gemm<..., float>)
  int64_t sA, lda;
  const double *B;  // NN: B[b][k][j] at B + b*sB + k*ldb + j;  NT: B[b][j][k] at ... j*ldb + k
  int64_t sB, ldb;
  double *O;        // dense: O[b][i][j] at O + b*sO + i*ldo + j
  int64_t sO, ldo;
  int m, n, k;
  double alpha;
  // packed-gradient epilogue (EPI == 1), only for i >= j:
  // G[b*ldg + tri0(i) + j] = (i == j ? 1 : 2) * (gv*V[b] * Pl[tri0(i)+j] + alpha*acc)
  double *G;
  int64_t ldg;
  const double *V;
  const double *Pl;
  double gv;
};

// grid = (tiles_i * tiles_j, batch).  O = alpha * A B (NN) or A B^T (NT).
// Double-buffered LDS with one barrier per kBK chunk: chunk c is multiplied
// out of buffer c&1 while chunk c+1 is stored from registers into the other
// buffer and chunk c+2 is fetched.  TA: element type of A (an FP32 A is
// widened to FP64 on load; the products and sums are FP64).
template <bool BT, int EPI, typename TA = double>
__global__ __launch_bounds__(kThreads) void k_gemm(GemmBatch g) {
  __shared__ double sA[2][kTM][kBK + 1];
  __shared__ double sB[2][kBK][kTM + 1];
  const int ntj = (g.n + kTM - 1) / kTM;
  const int ti = blockIdx.x / ntj, tj = blockIdx.x % ntj;
  if (EPI == 1 && ti < tj) return;  // strictly upper tile of a lower-triangle output
  const int64_t b = blockIdx.y;
  const int i0 = ti * kTM, j0 = tj * kTM;
  const TA *__restrict__ A;
  if constexpr (std::is_same<TA, float>::value) A = g.A32 + b * g.sA;
  else A = g.A + b * g.sA;
  const double *__restrict__ B = g.B + b * g.sB;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  v4d acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

  constexpr int PN = kTM * kBK / kThreads;
  static_assert(2 * PN <= 32, "element mask");
  double ra[PN], rb[PN];
  // Out-of-range elements load element 0 and are zeroed when stashed (bit it
  // of the mask for A, bit PN + it for B), so a chunk's loads stay in flight
  // across the multiply of the previous chunk.
  uint32_t okm = 0;
  auto fetch = [&](int k0) {
    okm = 0;
#pragma unroll
    for (int it = 0; it < PN; ++it) {
      const int e = tid + kThreads * it;
      const int r = e / kBK, kk = e % kBK;  // A: row r, column kk
      const int gi = i0 + r, gk = k0 + kk;
      const bool oka = gi < g.m && gk < g.k;
      ra[it] = double(A[oka ? int64_t(gi) * g.lda + gk : 0]);
      bool okb;
      if (BT) {  // B stored [j][k]: row r of the j tile
        const int gj = j0 + r;
        okb = gj < g.n && gk < g.k;
        rb[it] = B[okb ? int64_t(gj) * g.ldb + gk : 0];
      } else {  // B stored [k][j]
        const int kr = e / kTM, jj = e % kTM;
        const int gk2 = k0 + kr, gj = j0 + jj;
        okb = gk2 < g.k && gj < g.n;
        rb[it] = B[okb ? int64_t(gk2) * g.ldb + gj : 0];
      }
      okm |= (oka ? 1u << it : 0u) | (okb ? 1u << (PN + it) : 0u);
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int it = 0; it < PN; ++it) {
      const int e = tid + kThreads * it;
      const double va = ((okm >> it) & 1u) ? ra[it] : 0.0;
      const double vb = ((okm >> (PN + it)) & 1u) ? rb[it] : 0.0;
      sA[buf][e / kBK][e % kBK] = va;
      if (BT) sB[buf][e % kBK][e / kBK] = vb;
      else sB[buf][e / kTM][e % kTM] = vb;
    }
  };
  const int nch = (g.k + kBK - 1) / kBK;
  fetch(0);
  stash(0);
  if (nch > 1) fetch(kBK);
  __syncthreads();
  for (int c = 0; c < nch; ++c) {
    const int cur = c & 1;
#pragma unroll
    for (int ks = 0; ks < kBK / 4; ++ks) {
      const int kr = ks * 4 + (lane >> 4);
      double af[2], bf[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) af[i] = sA[cur][wm * 32 + i * 16 + (lane & 15)][kr];
#pragma unroll
      for (int j = 0; j < 2; ++j) bf[j] = sB[cur][kr][wn * 32 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = mfma64(af[i], bf[j], acc[i][j]);
    }
    if (c + 1 < nch) stash(cur ^ 1);
    if (c + 2 < nch) fetch((c + 2) * kBK);
    __syncthreads();
  }
  // result layout of v_mfma_f64_16x16x4f64: row 4*r + lane/16, column lane%16
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = i0 + wm * 32 + i * 16 + 4 * r + (lane >> 4);
        const int gj = j0 + wn * 32 + j * 16 + (lane & 15);
        if (gi >= g.m || gj >= g.n) continue;
        if (EPI == 0) {
          g.O[b * g.sO + int64_t(gi) * g.ldo + gj] = g.alpha * acc[i][j][r];
        } else {
          if (gi < gj) continue;
          const int64_t p = tri0(gi) + gj;
          const double s = (gi == gj) ? 1.0 : 2.0;
          g.G[b * g.ldg + p] = s * (g.gv * g.V[b] * g.Pl[p] + g.alpha * acc[i][j][r]);
        }
      }
}

// Dense symmetric n x n (row stride ld) from plain LT storage.
__global__ void k_unpack_lt(const double *__restrict__ lt, int n, double *__restrict__ out,
                            int64_t ld) {
  const int a = blockIdx.x;
  for (int b = threadIdx.x; b < n; b += blockDim.x) {
    const int hi = a > b ? a : b, lo = a > b ? b : a;
    out[int64_t(a) * ld + b] = lt[tri0(hi) + lo];
  }
}

// LT of scale * C C^T from the orbitals, C[a][k] = corb[k + a*o]; fold != 0
// doubles the off-diagonal (folded density), fold == 0 keeps plain storage.
__global__ void k_density_lt(const double *__restrict__ corb, int o, int n, double scale,
                             int fold, double *__restrict__ p) {
  const int a = blockIdx.x;
  for (int b = threadIdx.x; b <= a; b += blockDim.x) {
    double s = 0.0;
    for (int k = 0; k < o; ++k) s += corb[k + int64_t(a) * o] * corb[k + int64_t(b) * o];
    p[tri0(a) + b] = (fold && a != b ? 2.0 : 1.0) * scale * s;
  }
}

template <bool BT, int EPI, typename TA = double>
void gemm(Context *c, const GemmBatch &g, int batch) {
  const int tiles = ((g.m + kTM - 1) / kTM) * ((g.n + kTM - 1) / kTM);
  hipLaunchKernelGGL((k_gemm<BT, EPI, TA>), dim3(unsigned(tiles), unsigned(batch)), dim3(kThreads),
                     0, c->stream, g);
}

// dL[J] = s * (alpha V_J P + beta W[J] C^T) for the jn vectors whose W = C Z
// sits at W (n_pad x ldc per vector): lower triangle, packed, row stride ldd.
void grad_from_w(Context *c, int nocc, const double *W, int ldc, const double *v, const double *P,
                 double alpha, double beta, double *dl, int64_t ldd, int64_t jn) {
  GemmBatch g{};
  g.A = W; g.sA = int64_t(c->n_pad) * ldc; g.lda = ldc;
  g.B = static_cast<const double *>(c->w->cbank.ptr); g.sB = 0; g.ldb = ldc;  // NT: B[j][k] = C[b][k]
  g.m = c->n; g.n = c->n; g.k = nocc; g.alpha = beta;
  g.G = dl; g.ldg = ldd; g.V = v; g.Pl = P; g.gv = alpha;
  gemm<true, 1>(c, g, int(jn));
}

// P = 2 C C^T in plain LT storage (the misc workspace) unless the caller gave it.
int density_or(Context *c, int nocc, const double *corb, const double *p_lt, const double **P) {
  *P = p_lt;
  if (p_lt) return MJK_OK;
  int rc;
  if ((rc = ensure(c, c->w->misc, sizeof(double) * size_t(c->np)))) return rc;
  double *pw = static_cast<double *>(c->w->misc.ptr);
  hipLaunchKernelGGL(k_density_lt, dim3(unsigned(c->n)), dim3(256), 0, c->stream, corb, nocc, c->n,
                     2.0, 0, pw);
  *P = pw;
  return MJK_OK;
}

}  // namespace

void z_from_x(Context *c, const void *X, bool x_f32, int ldc, int nocc, int64_t jn, double *z) {
  GemmBatch g{};
  if (x_f32) g.A32 = static_cast<const float *>(X);
  else g.A = static_cast<const double *>(X);
  g.sA = int64_t(ldc) * c->n_pad; g.lda = c->n_pad;                          // X[J][k][a]
  g.B = static_cast<const double *>(c->w->cbank.ptr); g.sB = 0; g.ldb = ldc;  // C[a][l] (NN)
  g.O = z; g.sO = int64_t(nocc) * nocc; g.ldo = nocc;                         // Z[J][k][l]
  g.m = nocc; g.n = nocc; g.k = c->n; g.alpha = 1.0;
  if (x_f32) gemm<false, 0, float>(c, g, int(jn));
  else gemm<false, 0>(c, g, int(jn));
}

}  // namespace mjk

using namespace mjk;

extern "C" int mjk_density(mjk_context *c, int32_t nocc, const double *corb, double scale,
                           double *dlt) {
  if (!c) return MJK_ERR_ARG;
  if (nocc < 0 || (nocc > 0 && !corb) || !dlt) return set_error(c, MJK_ERR_ARG, "density: bad arguments");
  if (nocc == 0)
    return fill_zero(c, dlt, sizeof(double) * size_t(c->np));
  phase_begin(c);
  hipLaunchKernelGGL(k_density_lt, dim3(unsigned(c->n)), dim3(256), 0, c->stream, corb, nocc, c->n,
                     scale, 1, dlt);
  phase_end(c, MJK_PHASE_PACK);
  return hip_check(c, hipGetLastError(), "density launch");
}

extern "C" int mjk_grad_orbitals(mjk_context *c, int32_t nocc, const double *corb,
                                 const double *flt, double scale, double *dc) {
  if (!c) return MJK_ERR_ARG;
  if (nocc < 1 || !corb || !flt || !dc) return set_error(c, MJK_ERR_ARG, "grad_orbitals: bad arguments");
  const int n = c->n;
  int rc;
  if ((rc = ensure(c, c->w->misc, sizeof(double) * size_t(n) * n))) return rc;
  double *fd = static_cast<double *>(c->w->misc.ptr);
  phase_begin(c);
  hipLaunchKernelGGL(k_unpack_lt, dim3(unsigned(n)), dim3(256), 0, c->stream, flt, n, fd, int64_t(n));
  GemmBatch g{};
  g.A = fd; g.sA = 0; g.lda = n;           // F[a][b]
  g.B = corb; g.sB = 0; g.ldb = nocc;      // C[b][k] (NN)
  g.O = dc; g.sO = 0; g.ldo = nocc;        // dC[a][k]
  g.m = n; g.n = nocc; g.k = n; g.alpha = scale;
  gemm<false, 0>(c, g, 1);
  phase_end(c, MJK_PHASE_ASSEMBLY);
  return hip_check(c, hipGetLastError(), "grad_orbitals launch");
}

extern "C" int mjk_grad_vectors(mjk_context *c, int32_t nocc, const double *corb,
                                const double *p_lt, const double *v, double alpha, double beta,
                                double *dl, int64_t ldd) {
  if (!c) return MJK_ERR_ARG;
  if (nocc < 1 || nocc > 256 || !corb || !v || !dl || ldd < c->np)
    return set_error(c, MJK_ERR_ARG, "grad_vectors: bad arguments");
  if (c->nloaded != c->nchol)
    return set_error(c, MJK_ERR_SHORT_READ, "grad_vectors: not all Cholesky vectors were uploaded");
  const int n = c->n, n_pad = c->n_pad;
  const int64_t M = c->nchol, np = c->np;
  int rc;
  // bank with the orbitals (FP64), X chunk, Z and W per vector
  std::vector<int> off, opad;
  int ot = 0, tot = 0;
  const double *sets[1] = {corb};
  if ((rc = pack_bank(c, 1, &nocc, sets, off, opad, &ot, &tot))) return rc;
  const int ldc = 16 * ot;
  const double *cb = static_cast<const double *>(c->w->cbank.ptr);
  // per vector: X (ldc x n_pad) + Z (ldc x ldc) + W (n_pad x ldc)
  const int64_t per_vec = int64_t(ldc) * n_pad * 2 + int64_t(ldc) * ldc;
  int64_t jc = std::max<int64_t>(1, std::min<int64_t>(M, c->x_budget / (8 * per_vec)));
  const int64_t nchunk = ceil_div(M, jc);
  jc = ceil_div(M, nchunk);
  if ((rc = ensure(c, c->w->x, sizeof(double) * size_t(jc * per_vec)))) return rc;
  double *X = static_cast<double *>(c->w->x.ptr);
  double *Z = X + jc * ldc * int64_t(n_pad);
  double *W = Z + jc * ldc * int64_t(ldc);
  const double *P = nullptr;  // caller's plain LT, or 2 C C^T
  if ((rc = density_or(c, nocc, corb, p_lt, &P))) return rc;
  for (int64_t ch = 0; ch < nchunk; ++ch) {
    const int64_t j0 = ch * jc, jn = std::min(jc, M - j0);
    phase_begin(c);
    ht_launch(c, ot, cb, j0, jn, X);  // X[J][k][a] = (C^T L^J)[k][a]
    phase_end(c, MJK_PHASE_HALF_TRANSFORM);
    phase_begin(c);
    GemmBatch g{};
    // Z[J] = X[J] C   (o x o)
    g.A = X; g.sA = int64_t(ldc) * n_pad; g.lda = n_pad;
    g.B = cb; g.sB = 0; g.ldb = ldc;
    g.O = Z; g.sO = int64_t(ldc) * ldc; g.ldo = ldc;
    g.m = nocc; g.n = nocc; g.k = n; g.alpha = 1.0;
    gemm<false, 0>(c, g, int(jn));
    // W[J] = C Z[J]   (n x o)
    g = GemmBatch{};
    g.A = cb; g.sA = 0; g.lda = ldc;
    g.B = Z; g.sB = int64_t(ldc) * ldc; g.ldb = ldc;
    g.O = W; g.sO = int64_t(n_pad) * ldc; g.ldo = ldc;
    g.m = n; g.n = nocc; g.k = nocc; g.alpha = 1.0;
    gemm<false, 0>(c, g, int(jn));
    // dL[J] = s * (alpha V_J P + beta W[J] C^T), lower triangle, packed
    grad_from_w(c, nocc, W, ldc, v + j0, P, alpha, beta, dl + j0 * ldd, ldd, jn);
    phase_end(c, MJK_PHASE_BACKWARD);
  }
  return hip_check(c, hipGetLastError(), "grad_vectors launch");
}

extern "C" int mjk_grad_vectors_z(mjk_context *c, int32_t nocc, const double *corb,
                                  const double *p_lt, const double *v, const double *z,
                                  double alpha, double beta, double *dl, int64_t ldd) {
  if (!c) return MJK_ERR_ARG;
  if (nocc < 1 || nocc > 256 || !corb || !v || !z || !dl || ldd < c->np)
    return set_error(c, MJK_ERR_ARG, "grad_vectors_z: bad arguments");
  const int n_pad = c->n_pad;
  const int64_t M = c->nchol;
  int rc;
  std::vector<int> off, opad;
  int ot = 0, tot = 0;
  const double *sets[1] = {corb};
  if ((rc = pack_bank(c, 1, &nocc, sets, off, opad, &ot, &tot))) return rc;
  const int ldc = 16 * ot;
  const double *cb = static_cast<const double *>(c->w->cbank.ptr);
  // per vector: W = C Z (n_pad x ldc)
  const int64_t per_vec = int64_t(n_pad) * ldc;
  int64_t jc = std::max<int64_t>(1, std::min<int64_t>(M, c->x_budget / (8 * per_vec)));
  const int64_t nchunk = ceil_div(M, jc);
  jc = ceil_div(M, nchunk);
  if ((rc = ensure(c, c->w->x, sizeof(double) * size_t(jc * per_vec)))) return rc;
  double *W = static_cast<double *>(c->w->x.ptr);
  const double *P = nullptr;
  if ((rc = density_or(c, nocc, corb, p_lt, &P))) return rc;
  for (int64_t ch = 0; ch < nchunk; ++ch) {
    const int64_t j0 = ch * jc, jn = std::min(jc, M - j0);
    phase_begin(c);
    GemmBatch g{};
    // W[J] = C Z[J]   (n x o), Z dense nocc x nocc per vector
    g.A = cb; g.sA = 0; g.lda = ldc;
    g.B = z + j0 * int64_t(nocc) * nocc; g.sB = int64_t(nocc) * nocc; g.ldb = nocc;
    g.O = W; g.sO = per_vec; g.ldo = ldc;
    g.m = c->n; g.n = nocc; g.k = nocc; g.alpha = 1.0;
    gemm<false, 0>(c, g, int(jn));
    grad_from_w(c, nocc, W, ldc, v + j0, P, alpha, beta, dl + j0 * ldd, ldd, jn);
    phase_end(c, MJK_PHASE_BACKWARD);
  }
  return hip_check(c, hipGetLastError(), "grad_vectors_z launch");
}
