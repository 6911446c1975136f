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
  // launch shape, set by gemm(): output tiles per product (EPI == 1: the
  // lower-triangle tiles only) and tiles x batch work items
  int tiles;
  int64_t items;
};

// grid.x = round_up(tiles * batch, 8).  O = alpha * A B (NN) or A B^T (NT).
// Work item l = b * tiles + t (product b, output tile t); the block id is
// remapped so that consecutive items run on one XCD (the hardware deals
// blocks to the eight XCDs in turn): the tiles of one product then share
// that XCD's L2 for their common operand rows instead of each XCD fetching
// them from HBM.
// Double-buffered LDS with one barrier per kBK chunk: chunk c is multiplied
// out of buffer c&1 while chunk c+1 is stored from registers into the other
// buffer and chunk c+2 is fetched.  TA: element type of A (an FP32 A is
// widened to FP64 on load; the products and sums are FP64).
template <bool BT, int EPI, typename TA = double>
__global__ __launch_bounds__(kThreads) void k_gemm(GemmBatch g) {
  __shared__ double sA[2][kTM][kBK + 1];
  __shared__ double sB[2][kBK][kTM + 1];
  const int64_t nblk = gridDim.x;  // multiple of 8
  const int64_t l = (int64_t(blockIdx.x) & 7) * (nblk >> 3) + (int64_t(blockIdx.x) >> 3);
  if (l >= g.items) return;
  const int64_t b = l / g.tiles;
  const int t = int(l - b * g.tiles);
  int ti, tj;
  if constexpr (EPI == 1) {  // lower-triangle tile t = ti (ti + 1) / 2 + tj
    ti = int((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while (ti * (ti + 1) / 2 > t) --ti;
    while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
    tj = t - ti * (ti + 1) / 2;
  } else {
    const int ntj = (g.n + kTM - 1) / kTM;
    ti = t / ntj;
    tj = t - ti * ntj;
  }
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
  // Element it of a thread sits at a fixed (row, contraction column) of every
  // chunk, so its offset inside the product's operand is computed once and a
  // chunk adds only its first contraction index (32-bit byte offsets from the
  // product's base: every operand is below 2^28 elements, checked by gemm()).
  // Out-of-range elements load element 0 and are zeroed when stashed (bit it
  // of the mask for A, bit PN + it for B), so a chunk's loads stay in flight
  // across the multiply of the previous chunk.
  uint32_t aoff[PN], boff[PN];
  int akk[PN], bkk[PN];
  uint32_t rowok = 0;  // bit it: A row inside m; bit PN + it: B row / column inside n
#pragma unroll
  for (int it = 0; it < PN; ++it) {
    const int e = tid + kThreads * it;
    const int r = e / kBK, kk = e % kBK;  // A: row r, column kk
    const int gi = i0 + r;
    aoff[it] = uint32_t(gi) * uint32_t(g.lda) + uint32_t(kk);
    akk[it] = kk;
    rowok |= gi < g.m ? 1u << it : 0u;
    if (BT) {  // B stored [j][k]: row r of the j tile
      const int gj = j0 + r;
      boff[it] = uint32_t(gj) * uint32_t(g.ldb) + uint32_t(kk);
      bkk[it] = kk;
      rowok |= gj < g.n ? 1u << (PN + it) : 0u;
    } else {  // B stored [k][j]
      const int kr = e / kTM, jj = e % kTM, gj = j0 + jj;
      boff[it] = uint32_t(kr) * uint32_t(g.ldb) + uint32_t(gj);
      bkk[it] = kr;
      rowok |= gj < g.n ? 1u << (PN + it) : 0u;
    }
  }
  const char *__restrict__ Ab = reinterpret_cast<const char *>(A);
  const char *__restrict__ Bb = reinterpret_cast<const char *>(B);
  uint32_t okm = 0;
  auto fetch = [&](int k0) {
    okm = 0;
    const uint32_t bstep = BT ? uint32_t(k0) : uint32_t(k0) * uint32_t(g.ldb);
#pragma unroll
    for (int it = 0; it < PN; ++it) {
      const bool oka = ((rowok >> it) & 1u) && akk[it] + k0 < g.k;
      const bool okb = ((rowok >> (PN + it)) & 1u) && bkk[it] + k0 < g.k;
      ra[it] = double(*reinterpret_cast<const TA *>(
          Ab + (oka ? (aoff[it] + uint32_t(k0)) * uint32_t(sizeof(TA)) : 0u)));
      rb[it] = *reinterpret_cast<const double *>(Bb + (okb ? (boff[it] + bstep) * 8u : 0u));
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
    // operand fragments of the next 4-row step are read from LDS while the
    // current step's MFMAs run
    double af[2][2], bf[2][2];
    auto ld = [&](int ks, int q) {
      const int kr = ks * 4 + (lane >> 4);
#pragma unroll
      for (int i = 0; i < 2; ++i) af[q][i] = sA[cur][wm * 32 + i * 16 + (lane & 15)][kr];
#pragma unroll
      for (int j = 0; j < 2; ++j) bf[q][j] = sB[cur][kr][wn * 32 + j * 16 + (lane & 15)];
    };
    ld(0, 0);
#pragma unroll
    for (int ks = 0; ks < kBK / 4; ++ks) {
      const int q = ks & 1;
      if (ks + 1 < kBK / 4) ld(ks + 1, q ^ 1);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = mfma64(af[q][i], bf[q][j], acc[i][j]);
    }
    if (c + 1 < nch) stash(cur ^ 1);
    if (c + 2 < nch) fetch((c + 2) * kBK);
    __syncthreads();
  }
  // result layout of v_mfma_f64_16x16x4f64: row 4*r + lane/16, column lane%16
  if constexpr (EPI == 0) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gi = i0 + wm * 32 + i * 16 + 4 * r + (lane >> 4);
          const int gj = j0 + wn * 32 + j * 16 + (lane & 15);
          if (gi < g.m && gj < g.n) g.O[b * g.sO + int64_t(gi) * g.ldo + gj] = g.alpha * acc[i][j][r];
        }
  } else {
    // Every density value this lane needs is loaded before the first store:
    // with the loads and stores interleaved, each store would wait for its
    // load's round trip (G and P may alias as far as the compiler knows).
    const double *__restrict__ Pl = g.Pl;
    double *__restrict__ G = g.G + b * g.ldg;
    const double vb = g.gv * g.V[b];
    double pv[2][2][4];
    uint32_t pi[2][2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const uint32_t gi = uint32_t(i0 + wm * 32 + i * 16 + 4 * r + (lane >> 4));
          const uint32_t gj = uint32_t(j0 + wn * 32 + j * 16 + (lane & 15));
          const bool ok = gi < uint32_t(g.m) && gi >= gj;
          pi[i][j][r] = ok ? gi * (gi + 1u) / 2u + gj : ~0u;
          pv[i][j][r] = Pl[ok ? pi[i][j][r] : 0u];
        }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const uint32_t p = pi[i][j][r];
          if (p == ~0u) continue;
          const uint32_t gi = uint32_t(i0 + wm * 32 + i * 16 + 4 * r + (lane >> 4));
          const uint32_t gj = uint32_t(j0 + wn * 32 + j * 16 + (lane & 15));
          const double s = (gi == gj) ? 1.0 : 2.0;
          G[p] = s * (vb * pv[i][j][r] + g.alpha * acc[i][j][r]);
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
int gemm(Context *c, const GemmBatch &g0, int batch) {
  GemmBatch g = g0;
  // operands are addressed with 32-bit byte offsets inside one product
  constexpr int64_t kMaxElems = int64_t(1) << 28;
  if (int64_t(g.m) * g.lda >= kMaxElems || int64_t(g.k) * g.ldb >= kMaxElems ||
      int64_t(g.n) * g.ldb >= kMaxElems)
    return set_error(c, MJK_ERR_ARG, "gemm: operand of more than 2^28 elements");
  const int ti = (g.m + kTM - 1) / kTM, tj = (g.n + kTM - 1) / kTM;
  g.tiles = EPI == 1 ? ti * (ti + 1) / 2 : ti * tj;  // EPI == 1: square lower-triangle output
  g.items = int64_t(g.tiles) * batch;
  hipLaunchKernelGGL((k_gemm<BT, EPI, TA>), dim3(unsigned(round_up(g.items, 8))), dim3(kThreads), 0,
                     c->stream, g);
  return MJK_OK;
}

// dL[J] = s * (alpha V_J P + beta W[J] C^T) for the jn vectors whose W = C Z
// sits at W (n_pad x ldc per vector): lower triangle, packed, row stride ldd.
int grad_from_w(Context *c, int nocc, const double *W, int ldc, const double *v, const double *P,
                 double alpha, double beta, double *dl, int64_t ldd, int64_t jn) {
  GemmBatch g{};
  g.A = W; g.sA = int64_t(c->n_pad) * ldc; g.lda = ldc;
  g.B = static_cast<const double *>(c->w->cbank.ptr); g.sB = 0; g.ldb = ldc;  // NT: B[j][k] = C[b][k]
  g.m = c->n; g.n = c->n; g.k = nocc; g.alpha = beta;
  g.G = dl; g.ldg = ldd; g.V = v; g.Pl = P; g.gv = alpha;
  return gemm<true, 1>(c, g, int(jn));
}

// ---- fused vector gradient ------------------------------------------------
// G[J] (packed lower triangle, row stride ldg) = s_ab (gv V_J P_ab +
// beta (C Z^J C^T)_ab) for the vectors J whose Z^J (nocc x nocc) has element
// (k, l) at z + J sz + k ldz + l.
struct GradRows {
  const double *cb;  // orbital bank C[a][k]: n_pad rows of ldc, columns >= nocc zero
  const float *cb32; // the same bank in FP32 (FP32 products only)
  int ldc;
  const double *z;
  int64_t sz;
  int ldz;
  int o;             // nocc
  const double *v;   // V[J]
  const double *p;   // P, plain LT
  double gv, beta;
  double *g;
  int64_t ldg;
  int32_t rowbytes;  // bytes of one packed triangle n (n + 1) / 2 (< 2^31)
  int n, T;          // basis functions, row tiles of 64
  int64_t items;     // T * vectors
};
// Element index of a masked element in the packed triangle: byte offset
// 2^31, past every row, so the buffer load returns 0 and the store is dropped
// by the range check of the buffer resource (stride 0, rowbytes records).
constexpr uint32_t kMasked = 0x10000000u;
constexpr int kRsrcWord3 = 0x00020000;  // buffer resource word 3 of this architecture family
typedef int v2u __attribute__((ext_vector_type(2)));  // operand type of the 64-bit buffer builtins

typedef double f4d __attribute__((ext_vector_type(4)));
typedef float f4f __attribute__((ext_vector_type(4)));

// Arithmetic of k_grad_rows: FP64 MFMA, or FP32 MFMA (the split-precision
// backward's tail vectors: contractions of at most 128 terms, FP64 output).
// Operand layouts are those of v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32
// (A[l%16][l/16], B[l/16][l%16]); result row of register r: 4r + l/16 (FP64),
// 4(l/16) + r (FP32), column l%16.
struct GR64 {
  typedef double T;
  typedef v4d Acc;
  typedef f4d T4;
  __device__ static Acc mma(T a, T b, Acc c) { return mfma64(a, b, c); }
  __device__ static int row(int lane, int r) { return 4 * r + (lane >> 4); }
  __device__ static const T *bank(const GradRows &g) { return g.cb; }
};
struct GR32 {
  typedef float T;
  typedef v4f Acc;
  typedef f4f T4;
  __device__ static Acc mma(T a, T b, Acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
  __device__ static const T *bank(const GradRows &g) { return g.cb32; }
};

// Work item (vector J, row tile I of 64 AOs).  Phase 1 forms W_I = C_I Z^J (64
// x OC, OC = 16 OC16 >= nocc) with FP64 MFMA, both operands staged through LDS
// in 16-column chunks, double-buffered; W_I goes to LDS.  Phase 2 walks the
// column tiles j = 0..I: wave (wm, wn) forms its 32 x 32 quarter of W_I C_j^T
// with the A operand from LDS and the C_j fragments read straight from the
// bank, which every work item of the system reads and which stays in L2.  In a
// 16-column chunk lane group g takes the contraction indices 4g..4g+3 in the
// four MFMA steps (the A reads use the same order), so each fragment is one
// 32-byte load; the fragments of the chunk after next are in flight while a
// chunk is multiplied, and phase 2 needs no barrier.  Each finished block is
// written packed with the V P term, all P values loaded before the first
// store.  Within a vector the longest row tile runs first; consecutive items
// run on one XCD.  W never goes through HBM (the two-GEMM form writes n x ldc
// doubles per vector and reads them back once per column tile).
template <int OC16, class R>
__global__ __launch_bounds__(kThreads) void k_grad_rows(GradRows g) {
  typedef typename R::T T;
  typedef typename R::Acc Acc;
  constexpr int OC = 16 * OC16;
  constexpr bool F32 = sizeof(T) == 4;
  constexpr int WS = OC + 1;                                   // row stride of the phase-1 Z chunks
  // Row stride of W_I.  FP32: a lane's four A values of a chunk are one
  // 16-byte read; OC + 4 keeps rows 16-byte aligned and puts the 8 lanes of
  // a read pass (rows r..r+7, 32 banks of 4 bytes) on 8 distinct 16-byte
  // bank groups.  FP64 reads them one 8-byte value at a time (odd stride).
  constexpr int WW = F32 ? OC + 4 : OC + 1;
  constexpr int S1 = 2 * (kTM * (kBK + 1) + kBK * WS);         // phase-1 staging
  constexpr int SW = kTM * WW;
  __shared__ __attribute__((aligned(16))) T smem[S1 > SW ? S1 : SW];
  typedef T RowA[kTM][kBK + 1];
  typedef T RowB[kBK][WS];
  RowA *sA = reinterpret_cast<RowA *>(smem);                   // [2]
  RowB *sB = reinterpret_cast<RowB *>(smem + 2 * kTM * (kBK + 1));  // [2]
  T(*sW)[WW] = reinterpret_cast<T(*)[WW]>(smem);               // aliases the staging

  const int64_t nblk = gridDim.x;  // multiple of 8
  const int64_t l = (int64_t(blockIdx.x) & 7) * (nblk >> 3) + (int64_t(blockIdx.x) >> 3);
  if (l >= g.items) return;
  const int64_t J = l / g.T;
  const int I = g.T - 1 - int(l - J * g.T);
  const int i0 = I * kTM;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, gq = lane >> 4;
  const T *__restrict__ cb = R::bank(g);
  const int ldc = g.ldc, o = g.o;
  const double *__restrict__ zj = g.z + J * g.sz;
  const int ldz = g.ldz;

  // ---- phase 1: W_I = C_I Z^J; wave w owns rows 16w..16w+15, all OC columns
  Acc acc1[OC16];
#pragma unroll
  for (int t = 0; t < OC16; ++t) acc1[t] = Acc{0, 0, 0, 0};
  T ra[4], rb[OC16];
  uint32_t zok = 0;  // bit it: Z element it lies inside nocc x nocc (zeroed at stash time)
  auto fetch1 = [&](int k0) {
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int e = tid + kThreads * it, r = e / kBK, kk = e % kBK;
      ra[it] = cb[uint32_t(i0 + r) * uint32_t(ldc) + uint32_t(k0 + kk)];
    }
    zok = 0;
#pragma unroll
    for (int it = 0; it < OC16; ++it) {
      const int e = tid + kThreads * it, kr = e / OC, cc = e % OC, k = k0 + kr;
      const bool ok = k < o && cc < o;
      rb[it] = T(zj[ok ? uint32_t(k) * uint32_t(ldz) + uint32_t(cc) : 0u]);
      zok |= ok ? 1u << it : 0u;
    }
  };
  auto stash1 = [&](int buf) {
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int e = tid + kThreads * it;
      sA[buf][e / kBK][e % kBK] = ra[it];
    }
#pragma unroll
    for (int it = 0; it < OC16; ++it) {
      const int e = tid + kThreads * it;
      sB[buf][e / OC][e % OC] = ((zok >> it) & 1u) ? rb[it] : T(0);
    }
  };
  fetch1(0);
  stash1(0);
  if (OC16 > 1) fetch1(kBK);
  __syncthreads();
#pragma unroll 1
  for (int c = 0; c < OC16; ++c) {
    const int cur = c & 1;
#pragma unroll
    for (int ks = 0; ks < kBK / 4; ++ks) {
      const T a = sA[cur][16 * w + (lane & 15)][4 * ks + gq];
#pragma unroll
      for (int t = 0; t < OC16; ++t) acc1[t] = R::mma(a, sB[cur][4 * ks + gq][16 * t + (lane & 15)], acc1[t]);
    }
    if (c + 1 < OC16) stash1(cur ^ 1);
    if (c + 2 < OC16) fetch1((c + 2) * kBK);
    __syncthreads();
  }
  // every wave is past its last read of the staging buffers: W_I replaces them
#pragma unroll
  for (int t = 0; t < OC16; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) sW[16 * w + R::row(lane, r)][16 * t + (lane & 15)] = acc1[t][r];
  __syncthreads();

  // ---- phase 2: the blocks W_I C_j^T, j = 0..I, as one stream of chunks.
  // Every column tile is straight-line code (its OC16 chunks unrolled, the
  // fragment stage of each chunk a compile-time constant) issuing the same
  // global operations wherever it sits: the fragments of the chunk after
  // next are loaded even past the last chunk (clamped to it, never used),
  // and the epilogue stores every element, those above the diagonal or past
  // n with an offset past the buffer's range, which the hardware drops.  The vector memory counter of this
  // architecture counts loads and stores in one queue, so with a fixed count
  // per tile the compiler's waits are exact: the epilogue waits for the
  // tile's density values only, a chunk for its own fragments, and neither
  // for the stores of the tile before (conditional stores made it drain the
  // whole queue once per tile).
  const int wm = w >> 1, wn = w & 1;
  const int Q = (I + 1) * OC16;
  const double vj = g.gv * g.v[J];
  const double *__restrict__ P = g.p;
  double *__restrict__ G = g.g + J * g.ldg;
  const __amdgpu_buffer_rsrc_t rg = __builtin_amdgcn_make_buffer_rsrc(G, 0, g.rowbytes, kRsrcWord3);
  const __amdgpu_buffer_rsrc_t rp =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(P), 0, g.rowbytes, kRsrcWord3);
  typedef typename R::T4 T4;
  T4 bq[2][2];  // [stage][jj]: C_j fragments of the chunks in flight
  auto loadB = [&](int q, T4 (&dst)[2]) __attribute__((always_inline)) {
    q = q < Q ? q : Q - 1;
    const int j = q / OC16, kc = q - j * OC16;
#pragma unroll
    for (int jj = 0; jj < 2; ++jj) {
      const uint32_t row = uint32_t(j * kTM + wn * 32 + jj * 16 + (lane & 15));
      dst[jj] = *reinterpret_cast<const T4 *>(cb + row * uint32_t(ldc) + uint32_t(kc * kBK + 4 * gq));
    }
  };
  Acc acc2[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int jj = 0; jj < 2; ++jj) acc2[i][jj] = Acc{0, 0, 0, 0};
  // P values of the tile in progress, loaded at its start so the epilogue
  // only stores
  double pv[2][2][4];
  uint32_t pi[2][2][4];
  auto load_p = [&](int j) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int jj = 0; jj < 2; ++jj)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const uint32_t gi = uint32_t(i0 + wm * 32 + i * 16 + R::row(lane, r));
          const uint32_t gj = uint32_t(j * kTM + wn * 32 + jj * 16 + (lane & 15));
          const bool ok = gi < uint32_t(g.n) && gi >= gj;
          pi[i][jj][r] = (ok ? gi * (gi + 1u) / 2u + gj : kMasked) * 8u;  // byte offset
          pv[i][jj][r] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rp, int(pi[i][jj][r]), 0, 0));
        }
  };
  auto epilogue = [&](int j) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int jj = 0; jj < 2; ++jj)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const uint32_t gi = uint32_t(i0 + wm * 32 + i * 16 + R::row(lane, r));
          const uint32_t gj = uint32_t(j * kTM + wn * 32 + jj * 16 + (lane & 15));
          const double val = (gi == gj ? 1.0 : 2.0) * (vj * pv[i][jj][r] + g.beta * double(acc2[i][jj][r]));
          __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u, val), rg, int(pi[i][jj][r]), 0, 0);
        }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int jj = 0; jj < 2; ++jj) acc2[i][jj] = Acc{0, 0, 0, 0};
  };
  // Column tile j; PAR = parity of its first chunk in the stream (chunk q
  // uses fragment stage q & 1 and refills it with chunk q + 2).
  // The scheduler keeps each chunk to itself (sched_barrier): hoisting the
  // loads and LDS reads of later chunks raised the register count to the
  // point of one wave per SIMD.
  auto tile = [&](int j, auto parity) __attribute__((always_inline)) {
    constexpr int PAR = decltype(parity)::value;
    load_p(j);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int kc = 0; kc < OC16; ++kc) {
      const int S = (PAR + kc) & 1;
      const T4 b0 = bq[S][0], b1 = bq[S][1];
      loadB(j * OC16 + kc + 2, bq[S]);
      T4 a4[2];
      if constexpr (F32) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
          a4[i] = *reinterpret_cast<const T4 *>(&sW[wm * 32 + i * 16 + (lane & 15)][kc * kBK + 4 * gq]);
      }
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        T af[2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
          af[i] = F32 ? a4[i][ks] : sW[wm * 32 + i * 16 + (lane & 15)][kc * kBK + 4 * gq + ks];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          acc2[i][0] = R::mma(af[i], b0[ks], acc2[i][0]);
          acc2[i][1] = R::mma(af[i], b1[ks], acc2[i][1]);
        }
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    epilogue(j);
    __builtin_amdgcn_sched_barrier(0);
  };
  // tile j starts at chunk j OC16: parity 0 for even j, OC16 & 1 for odd j.
  // Tile 0 runs before the loop so that both ways into the loop end with an
  // epilogue and the wait counts at the loop head stay exact.
  loadB(0, bq[0]);
  loadB(1, bq[1]);
  tile(0, std::integral_constant<int, 0>{});
  int j = 1;
  for (; j + 1 <= I; j += 2) {
    tile(j, std::integral_constant<int, OC16 & 1>{});
    tile(j + 1, std::integral_constant<int, 0>{});
  }
  if (j <= I) tile(j, std::integral_constant<int, OC16 & 1>{});
}

template <int OC16>
void launch_grad_rows(Context *c, const GradRows &g) {
  if (g.cb32)
    hipLaunchKernelGGL((k_grad_rows<OC16, GR32>), dim3(unsigned(round_up(g.items, 8))), dim3(kThreads), 0,
                       c->stream, g);
  else
    hipLaunchKernelGGL((k_grad_rows<OC16, GR64>), dim3(unsigned(round_up(g.items, 8))), dim3(kThreads), 0,
                       c->stream, g);
}

// Whether the fused k_grad_rows takes this shape: nocc <= 128, 32-bit bank
// offsets, and a packed dL row under 2^31 bytes (the range of its buffer
// resources).  Other shapes form W = C Z first.
bool grad_rows_fits(const Context *c, int nocc, int ldc) {
  const int64_t rowbytes = int64_t(c->n) * (c->n + 1) / 2 * int64_t(sizeof(double));
  return (nocc + 15) / 16 <= 8 && int64_t(c->n_pad) * ldc < (int64_t(1) << 28) && rowbytes < (int64_t(1) << 31);
}

// dL of the jn vectors from Z in one fused launch; returns MJK_ERR_ARG when
// grad_rows_fits rejects the shape (the caller then forms W = C Z first).
// cb32: the bank in FP32 for FP32 products, or null for FP64.
int grad_rows(Context *c, int nocc, int ldc, const double *z, int64_t sz, int ldz, const double *v,
              const double *P, double alpha, double beta, double *dl, int64_t ldd, int64_t jn,
              const float *cb32) {
  const int oc16 = (nocc + 15) / 16;
  if (!grad_rows_fits(c, nocc, ldc)) return MJK_ERR_ARG;
  if (jn < 1) return MJK_OK;
  GradRows g{};
  g.rowbytes = int32_t(int64_t(c->n) * (c->n + 1) / 2 * int64_t(sizeof(double)));
  g.cb = static_cast<const double *>(c->w->cbank.ptr);
  g.cb32 = cb32;
  g.ldc = ldc;
  g.z = z;
  g.sz = sz;
  g.ldz = ldz;
  g.o = nocc;
  g.v = v;
  g.p = P;
  g.gv = alpha;
  g.beta = beta;
  g.g = dl;
  g.ldg = ldd;
  g.n = c->n;
  g.T = int(ceil_div(c->n, kTM));
  g.items = int64_t(g.T) * jn;
  switch (oc16) {
    case 1: launch_grad_rows<1>(c, g); break;
    case 2: launch_grad_rows<2>(c, g); break;
    case 3: launch_grad_rows<3>(c, g); break;
    case 4: launch_grad_rows<4>(c, g); break;
    case 5: launch_grad_rows<5>(c, g); break;
    case 6: launch_grad_rows<6>(c, g); break;
    case 7: launch_grad_rows<7>(c, g); break;
    default: launch_grad_rows<8>(c, g); break;
  }
  return MJK_OK;
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

// ---- Z^J = X^J C, lower blocks only ------------------------------------------
// Z^J = C^T L^J C is symmetric, so only its 16 x 16 blocks (i, j), j <= i are
// multiplied (OC16 = ceil(nocc/16) block rows, so at most 15 padding rows
// instead of a 64-wide tile's 63) and each off-diagonal block is also written
// transposed.  One workgroup per vector.  The block rows are dealt to the four
// waves at compile time (largest first, each to the wave with the fewest
// blocks), so every fragment index is static.  Per 16-row contraction chunk
// the bank rows C[a0 .. a0+15][0 .. 16 OC16) go through LDS (double-buffered,
// one barrier per chunk); a wave's X fragments are one 32-byte load per block
// row, lane group g taking the contraction indices 4g .. 4g+3 in the four
// MFMA steps (the B reads from LDS use the same order).
struct ZPlan {
  int nrows[4];
  int rows[4][16];
  int nblk[4];
};

constexpr ZPlan zplan(int oc16) {
  ZPlan p{};
  for (int i = oc16 - 1; i >= 0; --i) {  // row i holds i + 1 lower blocks
    int best = 0;
    for (int w = 1; w < 4; ++w)
      if (p.nblk[w] < p.nblk[best]) best = w;
    p.rows[best][p.nrows[best]++] = i;
    p.nblk[best] += i + 1;
  }
  return p;
}

template <int OC16, typename TX>
__global__ __launch_bounds__(kThreads) void k_zsym(const TX *__restrict__ X, int64_t sx, int ldx,
                                                  const double *__restrict__ cb, int ldc, int n,
                                                  int o, double *__restrict__ z, int64_t sz,
                                                  int ldz) {
  constexpr int OC = 16 * OC16;
  constexpr ZPlan P = zplan(OC16);
  constexpr int CL = (16 * OC) / kThreads > 0 ? (16 * OC) / kThreads : 1;  // bank doubles per thread
  __shared__ double sC[2][16 * OC];
  const int64_t J = blockIdx.x;
  const TX *__restrict__ xj = X + J * sx;
  const int tid = threadIdx.x, lane = tid & 63, gq = lane >> 4, w = tid >> 6;
  const int nch = (n + 15) / 16;
  double rc[CL];
  auto fetch = [&](int a0) {
#pragma unroll
    for (int q = 0; q < CL; ++q) {
      const int e = tid + kThreads * q;  // 16 * OC is a multiple of 256 or below it
      rc[q] = e < 16 * OC ? cb[(a0 + e / OC) * int64_t(ldc) + e % OC] : 0.0;
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int q = 0; q < CL; ++q) {
      const int e = tid + kThreads * q;
      if (e < 16 * OC) sC[buf][e] = rc[q];
    }
  };
  auto wave = [&](auto wc) {
    constexpr int W = decltype(wc)::value;
    constexpr int NR = P.nrows[W];
    v4d acc[NR > 0 ? NR : 1][OC16];
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
      for (int j = 0; j < OC16; ++j) acc[r][j] = v4d{0.0, 0.0, 0.0, 0.0};
    typedef TX TX4 __attribute__((ext_vector_type(4)));
    TX4 xs[2][NR > 0 ? NR : 1];  // [stage]: X fragments of chunk c in stage c & 1
    auto loadx = [&](int a0, TX4 (&dst)[NR > 0 ? NR : 1]) __attribute__((always_inline)) {
#pragma unroll
      for (int r = 0; r < NR; ++r)
        dst[r] = *reinterpret_cast<const TX4 *>(xj + int64_t(16 * P.rows[W][r] + (lane & 15)) * ldx + a0 +
                                                4 * gq);
    };
    // Every chunk issues the same loads (the ones past the last chunk clamped
    // to it, never used) and the X stages are compile-time (the loop runs two
    // chunks per trip), so no register copy of a load in flight and no
    // conditional load leaves the compiler a wait for the whole memory queue.
    auto chunk = [&](int c, auto stage) __attribute__((always_inline)) {
      constexpr int S = decltype(stage)::value;  // = c & 1: X stage and LDS buffer
      loadx(16 * (c + 1 < nch ? c + 1 : nch - 1), xs[S ^ 1]);
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        const double *__restrict__ brow = sC[S] + (4 * gq + ks) * OC + (lane & 15);
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          const double a = double(xs[S][r][ks]);
#pragma unroll
          for (int j = 0; j < OC16; ++j)
            if (j <= P.rows[W][r]) acc[r][j] = mfma64(a, brow[16 * j], acc[r][j]);
        }
      }
      stash(S ^ 1);
      fetch(16 * (c + 2 < nch ? c + 2 : nch - 1));
      __syncthreads();
    };
    fetch(0);
    stash(0);
    loadx(0, xs[0]);
    fetch(16 * (nch > 1 ? 1 : 0));
    __syncthreads();
    int c = 0;
    for (; c + 1 < nch; c += 2) {
      chunk(c, std::integral_constant<int, 0>{});
      chunk(c + 1, std::integral_constant<int, 1>{});
    }
    if (c < nch) chunk(c, std::integral_constant<int, 0>{});
    double *__restrict__ zj = z + J * sz;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int i = P.rows[W][r];
#pragma unroll
      for (int j = 0; j < OC16; ++j) {
        if (j > i) continue;
        const int l = 16 * j + (lane & 15);
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
          const int k = 16 * i + 4 * rr + gq;
          if (k < o && l < o) {
            zj[int64_t(k) * ldz + l] = acc[r][j][rr];
            if (j < i) zj[int64_t(l) * ldz + k] = acc[r][j][rr];
          }
        }
      }
    }
  };
  // every wave runs the chunk loop (its barriers) even when it owns no rows
  switch (w) {
    case 0: wave(std::integral_constant<int, 0>{}); break;
    case 1: wave(std::integral_constant<int, 1>{}); break;
    case 2: wave(std::integral_constant<int, 2>{}); break;
    default: wave(std::integral_constant<int, 3>{}); break;
  }
}

template <int OC16, typename TX>
void launch_zsym(Context *c, const TX *X, int64_t sx, const double *cb, int ldc, int nocc, double *z,
                 int64_t sz, int ldz, int64_t jn) {
  hipLaunchKernelGGL((k_zsym<OC16, TX>), dim3(unsigned(jn)), dim3(kThreads), 0, c->stream, X, sx,
                     c->n_pad, cb, ldc, c->n, nocc, z, sz, ldz);
}

// Z^J = X^J C for jn vectors (X rows of n_pad, sx apart per vector; bank cb of
// row stride ldc), Z^J at z + J sz with row stride ldz; nocc <= 128.
template <typename TX>
int zsym(Context *c, const TX *X, int64_t sx, const double *cb, int ldc, int nocc, double *z,
         int64_t sz, int ldz, int64_t jn) {
  if (nocc < 1 || nocc > 128 || jn < 1) return MJK_ERR_ARG;
  if (ldc < 16 * ((nocc + 15) / 16)) return set_error(c, MJK_ERR_ARG, "zsym: bank narrower than the blocks");
  switch ((nocc + 15) / 16) {
    case 1: launch_zsym<1>(c, X, sx, cb, ldc, nocc, z, sz, ldz, jn); break;
    case 2: launch_zsym<2>(c, X, sx, cb, ldc, nocc, z, sz, ldz, jn); break;
    case 3: launch_zsym<3>(c, X, sx, cb, ldc, nocc, z, sz, ldz, jn); break;
    case 4: launch_zsym<4>(c, X, sx, cb, ldc, nocc, z, sz, ldz, jn); break;
    case 5: launch_zsym<5>(c, X, sx, cb, ldc, nocc, z, sz, ldz, jn); break;
    case 6: launch_zsym<6>(c, X, sx, cb, ldc, nocc, z, sz, ldz, jn); break;
    case 7: launch_zsym<7>(c, X, sx, cb, ldc, nocc, z, sz, ldz, jn); break;
    default: launch_zsym<8>(c, X, sx, cb, ldc, nocc, z, sz, ldz, jn); break;
  }
  return hip_check(c, hipGetLastError(), "zsym launch");
}

}  // namespace

int z_from_x(Context *c, const void *X, bool x_f32, int ldc, int nocc, int64_t jn, double *z) {
  const double *cb = static_cast<const double *>(c->w->cbank.ptr);
  const int64_t sx = int64_t(ldc) * c->n_pad, sz = int64_t(nocc) * nocc;
  if (nocc <= 128)
    return x_f32 ? zsym(c, static_cast<const float *>(X), sx, cb, ldc, nocc, z, sz, nocc, jn)
                 : zsym(c, static_cast<const double *>(X), sx, cb, ldc, nocc, z, sz, nocc, jn);
  GemmBatch g{};
  if (x_f32) g.A32 = static_cast<const float *>(X);
  else g.A = static_cast<const double *>(X);
  g.sA = int64_t(ldc) * c->n_pad; g.lda = c->n_pad;                          // X[J][k][a]
  g.B = static_cast<const double *>(c->w->cbank.ptr); g.sB = 0; g.ldb = ldc;  // C[a][l] (NN)
  g.O = z; g.sO = int64_t(nocc) * nocc; g.ldo = nocc;                         // Z[J][k][l]
  g.m = nocc; g.n = nocc; g.k = c->n; g.alpha = 1.0;
  if (x_f32) return gemm<false, 0, float>(c, g, int(jn));
  return gemm<false, 0>(c, g, int(jn));
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
  if ((rc = gemm<false, 0>(c, g, 1))) return rc;
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
  const int64_t M = c->nchol;
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
    // Z[J] = X[J] C   (o x o), the same kernel as the forward's saved Z
    if (nocc <= 128) {
      if ((rc = zsym(c, static_cast<const double *>(X), int64_t(ldc) * n_pad, cb, ldc, nocc, Z,
                     int64_t(ldc) * ldc, ldc, jn)))
        return rc;
    } else {
      g.A = X; g.sA = int64_t(ldc) * n_pad; g.lda = n_pad;
      g.B = cb; g.sB = 0; g.ldb = ldc;
      g.O = Z; g.sO = int64_t(ldc) * ldc; g.ldo = ldc;
      g.m = nocc; g.n = nocc; g.k = n; g.alpha = 1.0;
      if ((rc = gemm<false, 0>(c, g, int(jn)))) return rc;
    }
    // dL[J] = s * (alpha V_J P + beta C Z[J] C^T), lower triangle, packed
    if (grad_rows(c, nocc, ldc, Z, int64_t(ldc) * ldc, ldc, v + j0, P, alpha, beta, dl + j0 * ldd,
                  ldd, jn, nullptr) != MJK_OK) {  // nocc > 128: W = C Z first
      g = GemmBatch{};
      g.A = cb; g.sA = 0; g.lda = ldc;
      g.B = Z; g.sB = int64_t(ldc) * ldc; g.ldb = ldc;
      g.O = W; g.sO = int64_t(n_pad) * ldc; g.ldo = ldc;
      g.m = n; g.n = nocc; g.k = nocc; g.alpha = 1.0;
      if ((rc = gemm<false, 0>(c, g, int(jn)))) return rc;
      if ((rc = grad_from_w(c, nocc, W, ldc, v + j0, P, alpha, beta, dl + j0 * ldd, ldd, jn))) return rc;
    }
    phase_end(c, MJK_PHASE_BACKWARD);
  }
  return hip_check(c, hipGetLastError(), "grad_vectors launch");
}

namespace {
__global__ void k_bank_f32(const double *__restrict__ src, float *__restrict__ dst, int64_t cnt) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < cnt; i += int64_t(gridDim.x) * blockDim.x)
    dst[i] = float(src[i]);
}

// dE/dL from a given Z; vectors [j_fp32, M) take FP32 products (nocc <= 128).
int grad_vectors_z_impl(Context *c, int32_t nocc, const double *corb, const double *p_lt,
                        const double *v, const double *z, double alpha, double beta, double *dl,
                        int64_t ldd, int64_t j_fp32, const char *who) {
  if (!c) return MJK_ERR_ARG;
  if (nocc < 1 || nocc > 256 || !corb || !v || !z || !dl || ldd < c->np)
    return set_error(c, MJK_ERR_ARG, std::string(who) + ": bad arguments");
  const int n_pad = c->n_pad;
  const int64_t M = c->nchol;
  if (j_fp32 < 0 || j_fp32 > M) return set_error(c, MJK_ERR_ARG, std::string(who) + ": split outside [0, nchol]");
  int rc;
  std::vector<int> off, opad;
  int ot = 0, tot = 0;
  const double *sets[1] = {corb};
  if ((rc = pack_bank(c, 1, &nocc, sets, off, opad, &ot, &tot))) return rc;
  const int ldc = 16 * ot;
  const double *cb = static_cast<const double *>(c->w->cbank.ptr);
  if (grad_rows_fits(c, nocc, ldc)) {  // one fused launch per precision: W = C Z stays on chip
    const double *P = nullptr;
    if ((rc = density_or(c, nocc, corb, p_lt, &P))) return rc;
    const float *cb32 = nullptr;
    if (j_fp32 < M) {
      const int64_t cnt = int64_t(n_pad) * ldc;
      if ((rc = ensure(c, c->w->gb32, sizeof(float) * size_t(cnt)))) return rc;
      hipLaunchKernelGGL(k_bank_f32, dim3(unsigned(std::min<int64_t>(ceil_div(cnt, 256), 4096))), dim3(256), 0,
                         c->stream, cb, static_cast<float *>(c->w->gb32.ptr), cnt);
      cb32 = static_cast<const float *>(c->w->gb32.ptr);
    }
    const int64_t sz = int64_t(nocc) * nocc;
    phase_begin(c);
    if ((rc = grad_rows(c, nocc, ldc, z, sz, nocc, v, P, alpha, beta, dl, ldd, j_fp32, nullptr)) ||
        (rc = grad_rows(c, nocc, ldc, z + j_fp32 * sz, sz, nocc, v + j_fp32, P, alpha, beta,
                        dl + j_fp32 * ldd, ldd, M - j_fp32, cb32)))
      return set_error(c, rc, std::string(who) + ": fused gradient launch");
    phase_end(c, MJK_PHASE_BACKWARD);
    return hip_check(c, hipGetLastError(), "grad_vectors_z launch");
  }
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
    if ((rc = gemm<false, 0>(c, g, int(jn)))) return rc;
    if ((rc = grad_from_w(c, nocc, W, ldc, v + j0, P, alpha, beta, dl + j0 * ldd, ldd, jn))) return rc;
    phase_end(c, MJK_PHASE_BACKWARD);
  }
  return hip_check(c, hipGetLastError(), "grad_vectors_z launch");
}
}  // namespace

extern "C" int mjk_grad_vectors_z(mjk_context *c, int32_t nocc, const double *corb,
                                  const double *p_lt, const double *v, const double *z,
                                  double alpha, double beta, double *dl, int64_t ldd) {
  if (!c) return MJK_ERR_ARG;
  return grad_vectors_z_impl(c, nocc, corb, p_lt, v, z, alpha, beta, dl, ldd, c->nchol, "grad_vectors_z");
}

extern "C" int mjk_grad_vectors_z_mixed(mjk_context *c, int32_t nocc, const double *corb,
                                        const double *p_lt, const double *v, const double *z,
                                        double alpha, double beta, double *dl, int64_t ldd,
                                        int64_t j_fp32) {
  if (!c) return MJK_ERR_ARG;
  return grad_vectors_z_impl(c, nocc, corb, p_lt, v, z, alpha, beta, dl, ldd, j_fp32, "grad_vectors_z_mixed");
}
