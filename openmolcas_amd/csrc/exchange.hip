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
#include <type_traits>

#include "mjk_internal.h"

namespace mjk {
namespace {

constexpr int kThreads = 256;

// Precision traits of the exchange kernels.  row(lane, r) is the row of the
// 16x16 result held in register r of `lane` (column = lane % 16); the FP64 and
// FP32 16x16x4 instructions of gfx950 differ here.  The A operand of lane l is
// A[l%16][l/16], the B operand B[l/16][l%16], for both.
struct F64 {
  typedef double T;
  typedef v4d Acc;
  static constexpr bool kFlush = false;  // accumulate in FP64 throughout
  __device__ static Acc mma(T a, T b, Acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lane, int r) { return 4 * r + (lane >> 4); }
};
struct F32 {
  typedef float T;
  typedef v4f Acc;
  static constexpr bool kFlush = true;  // FP32 partial sums folded into FP64
  __device__ static Acc mma(T a, T b, Acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// FP32 SYRK partial sums are folded into FP64 registers every kFlushRows
// contraction rows, so no FP32 sum runs over more than kFlushRows products.
constexpr int kFlushRows = 512;

// Densities of one SYRK launch: workgroup b belongs to the set d with
// item0[d] <= b < item0[d+1]; inside it, tile = local % ntiles and slice
// s = local / ntiles.  Set d contracts over bank columns off..off+opad-1 and
// writes partial tiles at kpart + kbase[d].
constexpr int kMaxSets = 16;
struct SyrkSets {
  int n;
  int off[kMaxSets];
  int opad[kMaxSets];
  int nsplit[kMaxSets];
  int item0[kMaxSets + 1];
  int64_t kbase[kMaxSets];
};

__device__ inline int find_set(const SyrkSets &sets, int b) {
  int d = 0;
  while (d + 1 < sets.n && b >= sets.item0[d + 1]) ++d;
  return d;
}

__device__ inline int64_t tri0(int64_t a) { return a * (a + 1) / 2; }
// Offset of packed row a within one vector in 32-bit arithmetic: contexts hold
// at most 46340 basis functions, so a(a+1) and the triangle fit in 32 bits.
__device__ inline uint32_t tri0u(uint32_t a) { return a * (a + 1u) / 2u; }

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
template <typename T>
__global__ void k_pack_orb(const double *__restrict__ corb, int nocc, int n, int ldc, int off,
                           T *__restrict__ cb) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= int64_t(nocc) * n) return;
  const int64_t a = i / nocc, k = i % nocc;
  cb[a * ldc + off + k] = T(corb[i]);
}

// FP32 shadow of the resident vectors for the mixed-precision exchange.
__global__ void k_to_f32(const double *__restrict__ src, float *__restrict__ dst, int64_t cnt) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < cnt;
       i += int64_t(gridDim.x) * blockDim.x)
    dst[i] = float(src[i]);
}

// ---- half-transform ------------------------------------------------------
// grid.x = round_up(ceil(jc/NJ) * ntile_a, 8), block = 256.  Workgroup (jr,
// a-tile of TA columns) computes, for the NJ vectors J = NJ*jr + q < jc,
// X[(J*LDC + k)*n_pad + a] = sum_b Cb[b*LDC + k] L^J_ab for all LDC = 16*OT
// bank columns k.  Wave w owns AO columns a0 + w*TA/4 .. a0 + (w+1)*TA/4 - 1
// (TA/64 MFMA column tiles) and all OT row tiles; each orbital fragment read
// from LDS feeds the MFMAs of all NJ vectors.  The next b-chunk is fetched
// into registers while the current one is multiplied out of LDS.
// FV: the workgroup also sums L^J_ba D_ba over the staged elements with b >= a
// (only "row-major" chunks hold such elements; each packed element is staged
// there by exactly one workgroup) into fv.vpart[tile * fv.m + j0 + J] -- the
// Coulomb pass V = L^T D without a read of L of its own, and the folded
// density read for half of the chunks only.
template <class P, int OT, int TA, int BK, int NJ, bool FV>
__global__ __launch_bounds__(kThreads) void k_ht(
    const typename P::T *__restrict__ L, int64_t ldl, int n, int n_pad,
    const typename P::T *__restrict__ Cb, int64_t j0, int64_t jc, int ntile_a, int64_t nblocks,
    typename P::T *__restrict__ X, const Screen sc, const FuseV fv) {
  typedef typename P::T T;
  typedef typename P::Acc Acc;
  static_assert(!FV || std::is_same<T, double>::value, "fused Coulomb pass is FP64 only");
  constexpr int LDC = 16 * OT;
  // Wave layout.  Narrow banks: wave w owns AO columns w*TA/4 .. (w+1)*TA/4-1
  // and all OT row tiles.  Wide banks (one vector, even OT >= 6, TA = 64): 2 x 2
  // waves, wave w owns row tiles (w>>1)*OT/2 .. and AO columns (w&1)*TA/2 ..,
  // so a step reads OT/2 + 2 fragments for OT MFMAs instead of OT + 1.
  constexpr int SPLIT = (NJ == 1 && OT % 2 == 0 && OT >= 6 && TA == 64) ? 2 : 1;
  constexpr int OTW = OT / SPLIT;         // row tiles per wave
  constexpr int NB = TA * SPLIT / 64;     // 16-column MFMA tiles per wave
  constexpr int SLD = TA + 1;            // LDS row stride of the L tile
  constexpr int CN = BK * LDC / kThreads;
  constexpr int LN = BK * TA / kThreads;
  static_assert(BK * LDC % kThreads == 0 && BK * TA % kThreads == 0, "tile shape");
  __shared__ __attribute__((aligned(16))) T sC[2][BK * LDC];
  __shared__ T sL[2][NJ][BK * SLD];
  const int64_t lb = xcd_logical(nblocks);
  if (lb < 0) return;
  const int64_t jr = lb / ntile_a;
  const int a0 = int(lb % ntile_a) * TA;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int t0 = SPLIT == 2 ? (w >> 1) * OTW : 0;
  const int c0 = SPLIT == 2 ? (w & 1) * (TA / 2) : w * (TA / 4);
  // vectors of this workgroup: present (J < jc) and not screened out
  bool live[NJ];
  const T *__restrict__ Lq[NJ];
#pragma unroll
  for (int q = 0; q < NJ; ++q) {
    const int64_t J = NJ * jr + q;
    live[q] = J < jc;
    Lq[q] = L + (j0 + (live[q] ? J : NJ * jr)) * ldl;
  }
  if (sc.lnorm) {  // screened: ||X tile||_F <= ||L^J[:, tile]||_F max_d ||C_d||_F
    bool any = false;
#pragma unroll
    for (int q = 0; q < NJ; ++q) {
      if (!live[q]) continue;
      const int64_t J = NJ * jr + q;
      double s2 = 0.0;
      for (int st = a0 / 64; st < (a0 + TA) / 64; ++st) {
        const double v = sc.lnorm[(j0 + J) * sc.nstrip + st];
        s2 += v * v;
      }
      if (sqrt(s2) * sc.cmax[0] < sc.thr_x) {
        T *__restrict__ xj = X + J * LDC * int64_t(n_pad);
        for (int e = tid; e < LDC * TA; e += kThreads) xj[int64_t(e / TA) * n_pad + a0 + e % TA] = T(0);
        if (tid == 0) atomicAdd(sc.cnt + 1, 1ull);
        live[q] = false;
      }
      any = any || live[q];
    }
    if (!any) return;
  }

  Acc acc[NJ][OTW][NB];
#pragma unroll
  for (int q = 0; q < NJ; ++q)
#pragma unroll
    for (int t = 0; t < OTW; ++t)
#pragma unroll
      for (int j = 0; j < NB; ++j) acc[q][t][j] = Acc{0, 0, 0, 0};
  T vp[NJ];  // FV: this thread's share of sum_ab L^J_ab P_ab
#pragma unroll
  for (int q = 0; q < NJ; ++q) vp[q] = T(0);

  if (a0 < n) {
    // orbital bank rows are contiguous: 16-byte pieces where the tile divides
    constexpr int VW = (CN * int(sizeof(T))) % 16 == 0 ? 16 / int(sizeof(T)) : 1;
    typedef T VT __attribute__((ext_vector_type(VW)));
    constexpr int CV = CN / VW;
    VT rc[CV];
    T rl[NJ][LN];
    T rd[FV ? LN : 1];  // FV: density elements of the staged L elements
    uint32_t vok = 0;   // FV: bit i: element i lies in the matrix with b >= a
    // bit i: element i lies inside the matrix.  The mask is applied when the
    // chunk is stashed, so the loads stay in flight across the multiply of
    // the current chunk (a select right after the load would wait for it).
    uint32_t lok = 0;
    static_assert(LN <= 32, "element mask");
    // Element i of a thread: "row-major" chunks (b > a and diagonal blocks)
    // walk e -> (bb = e / TA, aa = e % TA); "column" chunks (b < a, where the
    // packed row a holds b contiguous) walk e -> (bb = e % BK, aa = e / BK).
    // Chunks wholly inside the matrix and wholly below (column chunks) or above
    // the diagonal block (uniform per workgroup) step each thread's packed
    // offsets instead of forming tri0 per element: a thread's elements lie DA
    // columns (column chunks) or DB rows apart, and tri0(x + d) - tri0(x) =
    // d x + d (d + 1) / 2.  Diagonal and edge chunks take the general path.
    constexpr int DA = kThreads / BK, DB = kThreads / TA;
    static_assert(kThreads % BK == 0 && kThreads % TA == 0, "element walk");
    // The chunk kinds differ only in the packed index of each element: the
    // branches form the indices, and one load sequence after them fills rl
    // (and rd).  Loads issued inside the branches left their registers to be
    // merged by copies at the join, and a copy of a register whose load is in
    // flight waits for the whole memory queue: the chunk after next was then
    // waited for right after its loads were issued.
    auto fetch = [&](int b0) {
      const bool colmajor = b0 + BK <= a0;
      const bool inside = b0 + BK <= n && a0 + TA <= n;
      const T *__restrict__ cs = Cb + int64_t(b0) * LDC;
#pragma unroll
      for (int i = 0; i < CV; ++i)
        rc[i] = *reinterpret_cast<const VT *>(cs + (tid + kThreads * i) * VW);
      uint32_t idx[LN];
      if (inside && colmajor) {  // packed row a = a0 + tid/BK + DA i holds b0 + tid%BK
        uint32_t a = uint32_t(a0 + tid / BK);
        uint32_t x = tri0u(a) + uint32_t(b0 + tid % BK);
#pragma unroll
        for (int i = 0; i < LN; ++i) {
          idx[i] = x;
          x += uint32_t(DA) * a + uint32_t(DA * (DA + 1) / 2);
          a += uint32_t(DA);
        }
        lok = LN == 32 ? ~0u : (1u << LN) - 1u;
        if constexpr (FV) vok = 0u;
      } else if (inside && b0 >= a0 + TA) {  // packed row b = b0 + tid/TA + DB i at column a0 + tid%TA
        uint32_t b = uint32_t(b0 + tid / TA);
        uint32_t x = tri0u(b) + uint32_t(a0 + tid % TA);
#pragma unroll
        for (int i = 0; i < LN; ++i) {
          idx[i] = x;
          x += uint32_t(DB) * b + uint32_t(DB * (DB + 1) / 2);
          b += uint32_t(DB);
        }
        lok = LN == 32 ? ~0u : (1u << LN) - 1u;
        if constexpr (FV) vok = lok;
      } else {
#pragma unroll
        for (int i = 0; i < LN; ++i) {
          const int e = tid + kThreads * i;
          const int bb = colmajor ? e % BK : e / TA, aa = colmajor ? e / BK : e % TA;
          const int b = b0 + bb, a = a0 + aa;
          const bool ok = (b < n) && (a < n);
          const int hi = b >= a ? b : a, lo = b >= a ? a : b;
          idx[i] = ok ? tri0u(uint32_t(hi)) + uint32_t(lo) : 0u;  // branch-free masked load
          lok = ok ? (lok | (1u << i)) : (lok & ~(1u << i));
          if constexpr (FV) vok = (!colmajor && ok && b >= a) ? (vok | (1u << i)) : (vok & ~(1u << i));
        }
      }
#pragma unroll
      for (int i = 0; i < LN; ++i)
#pragma unroll
        for (int q = 0; q < NJ; ++q) rl[q][i] = Lq[q][idx[i]];
      if constexpr (FV) {  // column chunks hold no b >= a element: one cached address, masked by vok
#pragma unroll
        for (int i = 0; i < LN; ++i) rd[i] = fv.p[colmajor ? 0u : idx[i]];
      }
    };
    auto stash = [&](int b0, int buf) {
      const bool colmajor = b0 + BK <= a0;
#pragma unroll
      for (int i = 0; i < CV; ++i)
        *reinterpret_cast<VT *>(&sC[buf][(tid + kThreads * i) * VW]) = rc[i];
#pragma unroll
      for (int i = 0; i < LN; ++i) {
        const int e = tid + kThreads * i;
        const int bb = colmajor ? e % BK : e / TA, aa = colmajor ? e / BK : e % TA;
#pragma unroll
        for (int q = 0; q < NJ; ++q)
          sL[buf][q][bb * SLD + aa] = ((lok >> i) & 1u) ? rl[q][i] : T(0);
        if constexpr (FV) {
          if (!colmajor) {
            const T d = ((vok >> i) & 1u) ? rd[i] : T(0);
#pragma unroll
            for (int q = 0; q < NJ; ++q) vp[q] += rl[q][i] * d;
          }
        }
      }
    };
    // Double-buffered LDS: chunk c is multiplied out of buffer c&1 while
    // chunk c+1 is stored into the other buffer and chunk c+2 is fetched to
    // registers; one barrier per chunk.  The operand fragments of the next
    // 4-row step are read while the current step's MFMAs run.
    const int nch = (n + BK - 1) / BK;
    fetch(0);
    stash(0, 0);
    if (nch > 1) fetch(BK);
    __syncthreads();
    for (int c = 0; c < nch; ++c) {
      const int cur = c & 1;
      const T *__restrict__ sc = sC[cur];
      T af[2][OTW], bf[2][NJ][NB];
      auto ld = [&](int ks, int qq) {
        const int kr = ks * 4 + (lane >> 4);
#pragma unroll
        for (int q = 0; q < NJ; ++q)
#pragma unroll
          for (int j = 0; j < NB; ++j)
            bf[qq][q][j] = sL[cur][q][kr * SLD + c0 + j * 16 + (lane & 15)];
#pragma unroll
        for (int t = 0; t < OTW; ++t) af[qq][t] = sc[kr * LDC + 16 * (t0 + t) + (lane & 15)];
      };
      ld(0, 0);
#pragma unroll
      for (int ks = 0; ks < BK / 4; ++ks) {
        const int qq = ks & 1;
        if (ks + 1 < BK / 4) ld(ks + 1, qq ^ 1);
#pragma unroll
        for (int q = 0; q < NJ; ++q)
#pragma unroll
          for (int t = 0; t < OTW; ++t)
#pragma unroll
            for (int j = 0; j < NB; ++j) acc[q][t][j] = P::mma(af[qq][t], bf[qq][q][j], acc[q][t][j]);
      }
      if (c + 1 < nch) stash((c + 1) * BK, cur ^ 1);
      if (c + 2 < nch) fetch((c + 2) * BK);
      __syncthreads();
    }
  }
  if constexpr (FV) {  // workgroup sum in a fixed order: wave butterflies, then waves 0..3
    __shared__ double sV[NJ][kThreads / 64];
#pragma unroll
    for (int q = 0; q < NJ; ++q) {
      double v = vp[q];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
      if (lane == 0) sV[q][w] = v;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NJ; ++q)
      if (tid == q && live[q]) {
        double v = 0.0;
#pragma unroll
        for (int k = 0; k < kThreads / 64; ++k) v += sV[q][k];
        fv.vpart[int64_t(lb % ntile_a) * fv.m + j0 + NJ * jr + q] = v;
      }
  }
#pragma unroll
  for (int q = 0; q < NJ; ++q) {
    if (!live[q]) continue;
    T *__restrict__ xj = X + (NJ * jr + q) * LDC * int64_t(n_pad);
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const int a = a0 + c0 + j * 16 + (lane & 15);
#pragma unroll
      for (int t = 0; t < OTW; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          xj[int64_t(16 * (t0 + t) + P::row(lane, r)) * n_pad + a] = acc[q][t][j][r];
    }
  }
}

// ---- SYRK ------------------------------------------------------------------
// grid.x = ntiles * nsplit.  Tile t = A(A+1)/2 + B (A >= B) of TMxTM, split s
// covers chunk vectors [s*jper, min(jc,(s+1)*jper)).  The contraction runs
// over f = (J, k), k in [off, off+opad) of the bank columns.
// kpart[(s*ntiles + t)*TM*TM + al*TM + bl] (+)= sum_f X[f][A*TM+al] X[f][B*TM+bl]
// On diagonal tiles (A == B) only one operand is staged and the MFMA tiles
// strictly above the diagonal are skipped (those entries are never read).
template <class P, int TM, int BKF>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2))) void k_syrk(
    const typename P::T *__restrict__ X, int n_pad, int ldc, int64_t jc, int ntiles,
    const SyrkSets sets, double *__restrict__ kpart_all, int chunk, const Screen sc) {
  typedef typename P::T T;
  typedef typename P::Acc Acc;
  constexpr int WT = TM / 2;
  constexpr int MT = WT / 16;
  constexpr int VW = 16 / int(sizeof(T));  // elements per 16-byte load
  typedef T VT __attribute__((ext_vector_type(VW)));
  constexpr int PN = BKF * TM / kThreads / VW;
  static_assert(PN >= 1 && BKF * TM % (kThreads * VW) == 0, "tile shape");
  __shared__ __attribute__((aligned(16))) T sA[2][BKF * TM];
  __shared__ __attribute__((aligned(16))) T sB[2][BKF * TM];
  const int set = find_set(sets, blockIdx.x);
  const int local = blockIdx.x - sets.item0[set];
  const int tile = local % ntiles;
  const int s = local / ntiles;
  const int off = sets.off[set], opad = sets.opad[set];
  const int64_t jper = (jc + sets.nsplit[set] - 1) / sets.nsplit[set];
  double *__restrict__ kpart = kpart_all + sets.kbase[set];
  int A = int((sqrt(8.0 * tile + 1.0) - 1.0) * 0.5);
  while (A * (A + 1) / 2 > tile) --A;
  while ((A + 1) * (A + 2) / 2 <= tile) ++A;
  const int B = tile - A * (A + 1) / 2;
  const bool diag = (A == B);
  const int64_t ja = int64_t(s) * jper;
  const int64_t jz = min(jc, ja + jper);
  const int nf = int((jz > ja ? jz - ja : 0) * opad);  // < 2^31, checked on the host
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  const bool idle = diag && wm < wn;  // wave tile entirely above the diagonal
  // slot (chunk, slice): every chunk of vectors writes its own partial tiles
  double *__restrict__ kp =
      kpart + ((int64_t(chunk) * sets.nsplit[set] + s) * ntiles + tile) * TM * TM;
  if (sc.xn2 && nf > 0) {  // screened: |K_ab| of this slice <= sum_J ||X^J_A||_F ||X^J_B||_F
    double bound = 0.0;
    for (int64_t J = ja; J < jz; ++J) {
      const double *__restrict__ r = sc.xn2 + (int64_t(set) * jc + J) * sc.nstrip;
      double sa2 = 0.0, sb2 = 0.0;
#pragma unroll
      for (int st = 0; st < TM / 64; ++st) {
        sa2 += r[A * (TM / 64) + st];
        sb2 += r[B * (TM / 64) + st];
      }
      bound += sqrt(sa2 * sb2);
    }
    if (bound < sc.tau) {
      for (int e = tid; e < TM * TM; e += kThreads) kp[e] = 0.0;
      if (tid == 0) atomicAdd(sc.cnt, 1ull);
      return;
    }
  }

  Acc acc[MT][MT];
  double accd[P::kFlush ? MT : 1][P::kFlush ? MT : 1][4];
#pragma unroll
  for (int i = 0; i < MT; ++i)
#pragma unroll
    for (int j = 0; j < MT; ++j) acc[i][j] = Acc{0, 0, 0, 0};
  if constexpr (P::kFlush) {
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
      for (int j = 0; j < MT; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) accd[i][j][r] = 0.0;
  }
  auto flush = [&]() {
    if constexpr (P::kFlush) {
#pragma unroll
      for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < MT; ++j) {
#pragma unroll
          for (int r = 0; r < 4; ++r) accd[i][j][r] += double(acc[i][j][r]);
          acc[i][j] = Acc{0, 0, 0, 0};
        }
    }
  };

  VT ra[PN], rb[PN];
  uint32_t rok = 0;  // bit i: piece i lies inside the slice (applied at stash time)
  static_assert(PN <= 32, "piece mask");
  // Piece i of a thread (VW consecutive columns from cc) is row
  // f0 + ff0 + STEP*i; the row's (vector, bank column) pair is stepped
  // incrementally (one division per chunk).
  constexpr int STEP = kThreads * VW / TM;
  const int cc = (tid % (TM / VW)) * VW, ff0 = tid / (TM / VW);
  const int64_t row0 = (ja * ldc + off) * int64_t(n_pad);
  auto fetch = [&](int f0) {
    int f = f0 + ff0;
    int J = f / opad, kk = f - J * opad;
#pragma unroll
    for (int i = 0; i < PN; ++i) {
      const bool ok = f < nf;
      const int64_t ro = ok ? ((ja + J) * ldc + off + kk) * int64_t(n_pad) : row0;  // masked load
      ra[i] = *reinterpret_cast<const VT *>(X + ro + A * TM + cc);
      rb[i] = *reinterpret_cast<const VT *>(X + ro + B * TM + cc);
      rok = ok ? (rok | (1u << i)) : (rok & ~(1u << i));
      f += STEP;
      kk += STEP;
      while (kk >= opad) {  // opad >= 4: at most STEP / 4 wraps
        kk -= opad;
        ++J;
      }
    }
  };
  // FP32: a fragment read pass (32 lanes of 4 bytes, 32 banks) covers rows kr,
  // kr + 1 of TM floats, a multiple of 32 apart; column c of row f sits at
  // c ^ 16 (f & 1) so the two rows fall in different bank halves (the 16-byte
  // stores keep their four columns together)
  auto spos = [](int f, int c) {
    if constexpr (sizeof(T) == 4) return f * TM + (c ^ ((f & 1) << 4));
    else return f * TM + c;
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int i = 0; i < PN; ++i)
      *reinterpret_cast<VT *>(&sA[buf][spos(ff0 + STEP * i, cc)]) = ((rok >> i) & 1u) ? ra[i] : VT(0);
    if (!diag) {
#pragma unroll
      for (int i = 0; i < PN; ++i)
        *reinterpret_cast<VT *>(&sB[buf][spos(ff0 + STEP * i, cc)]) = ((rok >> i) & 1u) ? rb[i] : VT(0);
    }
  };
  // Double-buffered LDS: chunk c is multiplied out of buffer c&1 while chunk
  // c+1 is stored into the other buffer and chunk c+2 is fetched to
  // registers; one barrier per chunk.  Operand fragments of the next
  // 4-row step are read from LDS while the current step's MFMAs run.
  const int nch = (nf + BKF - 1) / BKF;
  if (nch > 0) {
    fetch(0);
    stash(0);
  }
  if (nch > 1) fetch(BKF);
  __syncthreads();
  for (int c = 0; c < nch; ++c) {
    const int cur = c & 1;
    if (!idle) {
      const T *__restrict__ sa = sA[cur];
      const T *__restrict__ sb = diag ? sA[cur] : sB[cur];
      T af[2][MT], bf[2][MT];
      auto ld = [&](int ks, int q) {
        const int kr = ks * 4 + (lane >> 4);
#pragma unroll
        for (int i = 0; i < MT; ++i) af[q][i] = sa[spos(kr, wm * WT + i * 16 + (lane & 15))];
#pragma unroll
        for (int j = 0; j < MT; ++j) bf[q][j] = sb[spos(kr, wn * WT + j * 16 + (lane & 15))];
      };
      // LO: a wave on the diagonal of a diagonal tile, MFMA tiles i >= j only
      // (a compile-time pattern, so the MFMAs issue back to back)
      auto steps = [&](auto lower_only) {
        constexpr bool LO = decltype(lower_only)::value;
        ld(0, 0);
#pragma unroll
        for (int ks = 0; ks < BKF / 4; ++ks) {
          const int q = ks & 1;
          if (ks + 1 < BKF / 4) ld(ks + 1, q ^ 1);
#pragma unroll
          for (int i = 0; i < MT; ++i)
#pragma unroll
            for (int j = 0; j < MT; ++j)
              if (!LO || i >= j) acc[i][j] = P::mma(af[q][i], bf[q][j], acc[i][j]);
        }
      };
      if (!diag || wm > wn) steps(std::false_type{});
      else steps(std::true_type{});
    }
    if constexpr (P::kFlush) {
      if (((c + 1) * BKF) % kFlushRows == 0) flush();
    }
    if (c + 1 < nch) stash(cur ^ 1);
    if (c + 2 < nch) fetch((c + 2) * BKF);
    __syncthreads();
  }
  flush();
#pragma unroll
  for (int i = 0; i < MT; ++i)
#pragma unroll
    for (int j = 0; j < MT; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int al = wm * WT + i * 16 + P::row(lane, r);
        const int bl = wn * WT + j * 16 + (lane & 15);
        double v;
        if constexpr (P::kFlush) v = accd[i][j][r];
        else v = acc[i][j][r];
        kp[al * TM + bl] = v;
      }
}

// ---- screening norms -------------------------------------------------------
__device__ inline double block_sum256(double x, double *red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
  if (lane == 0) red[wave] = x;
  __syncthreads();
  const double t = (red[0] + red[1]) + (red[2] + red[3]);
  __syncthreads();  // red is reused by the next call
  return t;
}

// out[J*nstrip + st] = ||L^J[:, 64*st .. 64*st+63]||_F of the symmetric matrix.
// grid = (M, nstrip), 256 threads.
__global__ __launch_bounds__(kThreads) void k_strip_norms(const double *__restrict__ L,
                                                          int64_t ldl, int n, int nstrip,
                                                          double *__restrict__ out) {
  __shared__ double red[4];
  const int64_t J = blockIdx.x;
  const int st = blockIdx.y;
  const double *__restrict__ Lj = L + J * ldl;
  double s = 0.0;
  for (int e = threadIdx.x; e < 64 * n; e += kThreads) {
    const int a = st * 64 + (e & 63), b = e >> 6;
    if (a < n) {
      const int hi = a > b ? a : b, lo = a > b ? b : a;
      const double v = Lj[tri0(hi) + lo];
      s += v * v;
    }
  }
  s = block_sum256(s, red);
  if (threadIdx.x == 0) out[J * nstrip + st] = sqrt(s);
}

// xn2[(d*jn + J)*nstrip + st] = sum over the rows of set d and the 64 AOs of
// strip st of X[J]^2.  grid = (jn, nstrip, nk), 256 threads.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_xnorms(const T *__restrict__ X, int ldc, int n_pad,
                                                     const SyrkSets sets, int nstrip,
                                                     double *__restrict__ xn2) {
  __shared__ double red[4];
  const int64_t J = blockIdx.x;
  const int st = blockIdx.y, d = blockIdx.z;
  const int off = sets.off[d], opad = sets.opad[d];
  double s = 0.0;
  for (int e = threadIdx.x; e < opad * 64; e += kThreads) {
    const double v = double(X[(J * ldc + off + (e >> 6)) * int64_t(n_pad) + st * 64 + (e & 63)]);
    s += v * v;
  }
  s = block_sum256(s, red);
  if (threadIdx.x == 0) xn2[(int64_t(d) * gridDim.x + J) * nstrip + st] = s;
}

// out[d] = ||C_d||_F from the bank (rows a < n, columns off_d .. off_d+opad_d-1),
// out[nk] = max_d out[d].  One workgroup.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_cnorm(const T *__restrict__ Cb, int ldc, int n,
                                                    const SyrkSets sets, double *__restrict__ out) {
  __shared__ double red[4];
  double mx = 0.0;
  for (int d = 0; d < sets.n; ++d) {
    const int off = sets.off[d], opad = sets.opad[d];
    double s = 0.0;
    for (int e = threadIdx.x; e < n * opad; e += kThreads) {
      const double v = double(Cb[int64_t(e / opad) * ldc + off + e % opad]);
      s += v * v;
    }
    s = block_sum256(s, red);
    mx = fmax(mx, sqrt(s));
    if (threadIdx.x == 0) out[d] = sqrt(s);
  }
  if (threadIdx.x == 0) out[sets.n] = mx;
}

// ---- split-BF16, three products ---------------------------------------------
// x = hi + lo with hi = bf16(x), lo = bf16(x - hi); a product is taken as
// lo*hi + hi*lo + hi*hi (lo*lo dropped) with v_mfma_f32_16x16x32_bf16 and FP32
// accumulation.  That MFMA takes 8 consecutive contraction indices per lane
// (lane l: row/column l%16, indices 8*(l/16)..8*(l/16)+7), so both operands are
// staged contraction-contiguous in LDS: rows of kB3S bf16 (32 used, padded to
// keep 16-byte alignment).  Result layout: row 4*(l/16) + r, column l%16.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kB3K = 32;          // contraction indices per MFMA
constexpr int kB3S = kB3K + 8;    // LDS row stride in bf16 elements

// LDS position of contraction index k (0..31) of row r in a split-BF16 tile:
// the 16-byte granule k / 8 is XOR-ed with bits 3-4 of the row.  A fragment
// read pass (8 lanes of 16 bytes, 32 banks of 4 bytes) covers 8 consecutive
// rows of one granule and stays conflict-free (rows 80 bytes apart); stores
// whose lanes walk rows 4 or 8 apart (SYRK staging, the half-transform's row
// chunks) no longer pile onto two bank groups.
__device__ __forceinline__ int b3pos(int r, int k) {
  return r * kB3S + ((((k >> 3) ^ (r >> 3)) & 3) << 3) + (k & 7);
}

__device__ inline void split_bf16(float x, __bf16 &hi, __bf16 &lo) {
  hi = __bf16(x);
  lo = __bf16(x - float(hi));
}

__device__ inline v4f mma_b3(bf16x8 ah, bf16x8 al, bf16x8 bh, bf16x8 bl, v4f acc) {
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc, 0, 0, 0);
}

// Transposed split bank: ch/cl[(off + k)*n_pad + a] = hi/lo(corb[k + a*nocc]).
__global__ void k_pack_orb_b3(const double *__restrict__ corb, int nocc, int n, int n_pad, int off,
                              __bf16 *__restrict__ ch, __bf16 *__restrict__ cl) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= int64_t(nocc) * n) return;
  const int64_t a = i / nocc, k = i % nocc;
  __bf16 h, l;
  split_bf16(float(corb[i]), h, l);
  ch[(off + k) * int64_t(n_pad) + a] = h;
  cl[(off + k) * int64_t(n_pad) + a] = l;
}

// Half-transform, same work decomposition as k_ht: X[(jr*LDC + k)*n_pad + a].
template <int OT, int TA>
__global__ __launch_bounds__(kThreads) void k_ht_b3(
    const float *__restrict__ L, int64_t ldl, int n, int n_pad, const __bf16 *__restrict__ CbH,
    const __bf16 *__restrict__ CbL, int64_t j0, int ntile_a, int64_t nblocks,
    float *__restrict__ X) {
  constexpr int LDC = 16 * OT;
  constexpr int NB = TA / 64;
  constexpr int LN = kB3K * TA / kThreads;
  constexpr int CCH = LDC * kB3K / 8;  // 16-byte pieces of one bank array per chunk
  constexpr int CN = (CCH + kThreads - 1) / kThreads;
  static_assert(kB3K * TA % kThreads == 0, "tile shape");
  __shared__ __attribute__((aligned(16))) __bf16 sCh[LDC * kB3S];
  __shared__ __attribute__((aligned(16))) __bf16 sCl[LDC * kB3S];
  __shared__ __attribute__((aligned(16))) __bf16 sLh[TA * kB3S];
  __shared__ __attribute__((aligned(16))) __bf16 sLl[TA * kB3S];
  const int64_t lb = xcd_logical(nblocks);
  if (lb < 0) return;
  const int64_t jr = lb / ntile_a;
  const int a0 = int(lb % ntile_a) * TA;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const float *__restrict__ Lj = L + (j0 + jr) * ldl;

  v4f acc[OT][NB];
#pragma unroll
  for (int t = 0; t < OT; ++t)
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[t][j] = v4f{0, 0, 0, 0};

  if (a0 < n) {
    // L staging: thread t owns LG groups of four consecutive contraction
    // indices b (one AO column a each) and writes each group as one 8-byte LDS
    // store per hi/lo array.  Row chunks (packed row b holds a contiguous) put
    // a on consecutive lanes; column chunks (packed row a holds b contiguous)
    // put the eight groups of one column on consecutive lanes.
    typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
    constexpr int LG = LN / 4;
    static_assert(LN % 4 == 0, "staging shape");
    float rl[LG][4];
    int lo[LG];
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));  // 16 bytes, promotable to registers
    u32x4 rch[CN], rcl[CN];
    auto fetch = [&](int b0) {
#pragma unroll
      for (int i = 0; i < CN; ++i) {
        const int p = tid + kThreads * i;
        if (p < CCH) {
          const int64_t src = int64_t(p / 4) * n_pad + b0 + (p % 4) * 8;
          rch[i] = *reinterpret_cast<const u32x4 *>(CbH + src);
          rcl[i] = *reinterpret_cast<const u32x4 *>(CbL + src);
        }
      }
      // chunks wholly inside the matrix step each thread's packed offsets
      // (tri0(x + d) - tri0(x) = d x + d (d + 1) / 2) instead of forming tri0
      // per element; diagonal and edge chunks take the general path
      const bool inside = b0 + kB3K <= n && a0 + TA <= n;
      constexpr int GA = kThreads / (kB3K / 4);  // column step between a thread's groups
      constexpr int GB = 4 * (kThreads / TA);     // row step between a thread's groups
      if (inside && b0 + kB3K <= a0) {  // packed row a holds b0 + 4q .. contiguous
        uint32_t a = uint32_t(a0 + tid / (kB3K / 4));
        uint32_t off = tri0u(a) + uint32_t(b0 + 4 * (tid % (kB3K / 4)));
#pragma unroll
        for (int i = 0; i < LG; ++i) {
#pragma unroll
          for (int k = 0; k < 4; ++k) rl[i][k] = Lj[off + uint32_t(k)];
          lo[i] = b3pos(tid / (kB3K / 4) + GA * i, 4 * (tid % (kB3K / 4)));
          off += uint32_t(GA) * a + uint32_t(GA * (GA + 1) / 2);
          a += uint32_t(GA);
        }
      } else if (inside && b0 >= a0 + TA) {  // packed rows b, b+1, b+2, b+3 at column a
        const uint32_t a = uint32_t(a0 + tid % TA);
        uint32_t b = uint32_t(b0 + 4 * (tid / TA));
#pragma unroll
        for (int i = 0; i < LG; ++i) {
          uint32_t idx = tri0u(b) + a;
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            rl[i][k] = Lj[idx];
            idx += b + uint32_t(k) + 1u;
          }
          lo[i] = b3pos(tid % TA, 4 * (tid / TA + (GB / 4) * i));
          b += uint32_t(GB);
        }
      } else if (b0 + kB3K <= a0) {  // every b < a
#pragma unroll
        for (int i = 0; i < LG; ++i) {
          const int g = tid + kThreads * i, q = g % (kB3K / 4), aa = g / (kB3K / 4), a = a0 + aa;
          const float *__restrict__ row = Lj + (tri0u(uint32_t(a)) + uint32_t(b0 + 4 * q));
#pragma unroll
          for (int k = 0; k < 4; ++k) rl[i][k] = (a < n) ? row[k] : 0.f;
          lo[i] = b3pos(aa, 4 * q);
        }
      } else {
#pragma unroll
        for (int i = 0; i < LG; ++i) {
          const int g = tid + kThreads * i, aa = g % TA, q = g / TA, a = a0 + aa;
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const int b = b0 + 4 * q + k;
            float v = 0.f;
            if (b < n && a < n)
              v = (b >= a) ? Lj[tri0u(uint32_t(b)) + uint32_t(a)] : Lj[tri0u(uint32_t(a)) + uint32_t(b)];
            rl[i][k] = v;
          }
          lo[i] = b3pos(aa, 4 * q);
        }
      }
    };
    fetch(0);
    for (int b0 = 0; b0 < n; b0 += kB3K) {
      if (b0) __syncthreads();
#pragma unroll
      for (int i = 0; i < CN; ++i) {
        const int p = tid + kThreads * i;
        if (p < CCH) {
          const int dst = (p / 4) * kB3S + (p % 4) * 8;
          *reinterpret_cast<u32x4 *>(sCh + dst) = rch[i];
          *reinterpret_cast<u32x4 *>(sCl + dst) = rcl[i];
        }
      }
#pragma unroll
      for (int i = 0; i < LG; ++i) {
        bf16x4 h, l;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          __bf16 hh, ll;
          split_bf16(rl[i][k], hh, ll);
          h[k] = hh;
          l[k] = ll;
        }
        *reinterpret_cast<bf16x4 *>(sLh + lo[i]) = h;
        *reinterpret_cast<bf16x4 *>(sLl + lo[i]) = l;
      }
      __syncthreads();
      if (b0 + kB3K < n) fetch(b0 + kB3K);
      const int ko = 8 * (lane >> 4);
      bf16x8 bh[NB], bl[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        const int r = w * (TA / 4) + 16 * j + (lane & 15);
        bh[j] = *reinterpret_cast<const bf16x8 *>(sLh + b3pos(r, ko));
        bl[j] = *reinterpret_cast<const bf16x8 *>(sLl + b3pos(r, ko));
      }
#pragma unroll
      for (int t = 0; t < OT; ++t) {
        const int r = 16 * t + (lane & 15);
        const bf16x8 ah = *reinterpret_cast<const bf16x8 *>(sCh + r * kB3S + ko);
        const bf16x8 al = *reinterpret_cast<const bf16x8 *>(sCl + r * kB3S + ko);
#pragma unroll
        for (int j = 0; j < NB; ++j) acc[t][j] = mma_b3(ah, al, bh[j], bl[j], acc[t][j]);
      }
    }
  }
  float *__restrict__ xj = X + jr * LDC * int64_t(n_pad);
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    const int a = a0 + w * (TA / 4) + j * 16 + (lane & 15);
#pragma unroll
    for (int t = 0; t < OT; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) xj[int64_t(16 * t + 4 * (lane >> 4) + r) * n_pad + a] = acc[t][j][r];
  }
}

// SYRK on FP32 X with split-BF16 products; same tiles, slices and FP64 flush
// as k_syrk.
template <int TM>
__global__ __launch_bounds__(kThreads) void k_syrk_b3(
    const float *__restrict__ X, int n_pad, int ldc, int64_t jc, int ntiles, const SyrkSets sets,
    double *__restrict__ kpart_all, int chunk) {
  constexpr int WT = TM / 2;
  constexpr int MT = WT / 16;
  __shared__ __attribute__((aligned(16))) __bf16 sAh[TM * kB3S];
  __shared__ __attribute__((aligned(16))) __bf16 sAl[TM * kB3S];
  __shared__ __attribute__((aligned(16))) __bf16 sBh[TM * kB3S];
  __shared__ __attribute__((aligned(16))) __bf16 sBl[TM * kB3S];
  const int set = find_set(sets, blockIdx.x);
  const int local = blockIdx.x - sets.item0[set];
  const int tile = local % ntiles;
  const int s = local / ntiles;
  const int off = sets.off[set], opad = sets.opad[set];
  const int64_t jper = (jc + sets.nsplit[set] - 1) / sets.nsplit[set];
  double *__restrict__ kpart = kpart_all + sets.kbase[set];
  int A = int((sqrt(8.0 * tile + 1.0) - 1.0) * 0.5);
  while (A * (A + 1) / 2 > tile) --A;
  while ((A + 1) * (A + 2) / 2 <= tile) ++A;
  const int B = tile - A * (A + 1) / 2;
  const bool diag = (A == B);
  const int64_t ja = int64_t(s) * jper;
  const int64_t jz = min(jc, ja + jper);
  const int nf = int((jz > ja ? jz - ja : 0) * opad);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  const bool idle = diag && wm < wn;

  v4f acc[MT][MT];
  double accd[MT][MT][4];
#pragma unroll
  for (int i = 0; i < MT; ++i)
#pragma unroll
    for (int j = 0; j < MT; ++j) {
      acc[i][j] = v4f{0, 0, 0, 0};
#pragma unroll
      for (int r = 0; r < 4; ++r) accd[i][j][r] = 0.0;
    }
  // Thread t stages columns 4cq..4cq+3 of rows RQ*fq .. RQ*fq+RQ-1 of each
  // chunk (cq = t % (TM/4), fq = t / (TM/4)): RQ 16-byte row loads per operand
  // and, per column, one LDS write of RQ consecutive contraction indices for
  // each of hi and lo.  The RQ rows share one vector J (chunk starts, RQ*fq
  // and opad are multiples of RQ, which divides 4).
  constexpr int RQ = kB3K * TM / (4 * kThreads);
  static_assert(RQ == 2 || RQ == 4, "staging shape");
  typedef float f4 __attribute__((ext_vector_type(4)));
  typedef __bf16 bfq __attribute__((ext_vector_type(RQ)));
  const int cq = tid % (TM / 4), fq = tid / (TM / 4);
  f4 ra[RQ], rb[RQ];
  auto fetch = [&](int f0) {
    const int f = f0 + RQ * fq;
    const int J = f / opad, kk = f - J * opad;
    const bool ok = f < nf;  // nf is a multiple of 4: the group is all in or all out
    const int64_t ro = ok ? ((ja + J) * ldc + off + kk) * int64_t(n_pad) : 0;
#pragma unroll
    for (int k = 0; k < RQ; ++k) {
      f4 va = f4(0.f), vb = f4(0.f);
      if (ok) {
        va = *reinterpret_cast<const f4 *>(X + ro + k * int64_t(n_pad) + A * TM + 4 * cq);
        if (!diag) vb = *reinterpret_cast<const f4 *>(X + ro + k * int64_t(n_pad) + B * TM + 4 * cq);
      }
      ra[k] = va;
      rb[k] = vb;
    }
  };
  auto stash_one = [&](const f4 (&r)[RQ], __bf16 *sh, __bf16 *sl) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      bfq h, l;
#pragma unroll
      for (int k = 0; k < RQ; ++k) {
        __bf16 hh, ll;
        split_bf16(r[k][c], hh, ll);
        h[k] = hh;
        l[k] = ll;
      }
      *reinterpret_cast<bfq *>(sh + b3pos(4 * cq + c, RQ * fq)) = h;
      *reinterpret_cast<bfq *>(sl + b3pos(4 * cq + c, RQ * fq)) = l;
    }
  };
  if (nf > 0) fetch(0);
  for (int f0 = 0; f0 < nf; f0 += kB3K) {
    if (f0) __syncthreads();
    stash_one(ra, sAh, sAl);
    if (!diag) stash_one(rb, sBh, sBl);
    __syncthreads();
    if (f0 + kB3K < nf) fetch(f0 + kB3K);
    if (!idle) {
      const __bf16 *sbh = diag ? sAh : sBh;
      const __bf16 *sbl = diag ? sAl : sBl;
      const int ko = 8 * (lane >> 4);
      bf16x8 ah[MT], al[MT], bh[MT], bl[MT];
#pragma unroll
      for (int i = 0; i < MT; ++i) {
        const int r = wm * WT + 16 * i + (lane & 15);
        ah[i] = *reinterpret_cast<const bf16x8 *>(sAh + b3pos(r, ko));
        al[i] = *reinterpret_cast<const bf16x8 *>(sAl + b3pos(r, ko));
      }
#pragma unroll
      for (int j = 0; j < MT; ++j) {
        const int r = wn * WT + 16 * j + (lane & 15);
        bh[j] = *reinterpret_cast<const bf16x8 *>(sbh + b3pos(r, ko));
        bl[j] = *reinterpret_cast<const bf16x8 *>(sbl + b3pos(r, ko));
      }
      if (!diag || wm > wn) {
#pragma unroll
        for (int i = 0; i < MT; ++i)
#pragma unroll
          for (int j = 0; j < MT; ++j) acc[i][j] = mma_b3(ah[i], al[i], bh[j], bl[j], acc[i][j]);
      } else {  // diagonal wave of a diagonal tile: MFMA tiles i >= j
#pragma unroll
        for (int i = 0; i < MT; ++i)
#pragma unroll
          for (int j = 0; j <= i; ++j) acc[i][j] = mma_b3(ah[i], al[i], bh[j], bl[j], acc[i][j]);
      }
    }
    if ((f0 + kB3K) % kFlushRows == 0) {
#pragma unroll
      for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < MT; ++j) {
#pragma unroll
          for (int r = 0; r < 4; ++r) accd[i][j][r] += double(acc[i][j][r]);
          acc[i][j] = v4f{0, 0, 0, 0};
        }
    }
  }
  // slot (chunk, slice): every chunk of vectors writes its own partial tiles
  double *__restrict__ kp =
      kpart + ((int64_t(chunk) * sets.nsplit[set] + s) * ntiles + tile) * TM * TM;
#pragma unroll
  for (int i = 0; i < MT; ++i)
#pragma unroll
    for (int j = 0; j < MT; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int al_ = wm * WT + i * 16 + 4 * (lane >> 4) + r;
        const int bl_ = wn * WT + j * 16 + (lane & 15);
        kp[al_ * TM + bl_] = accd[i][j][r] + double(acc[i][j][r]);
      }
}

// klt[d*ldk + a*(a+1)/2 + b] = sum_s kpart[kbase[d] + (s*ntiles + tile(a,b))*TM*TM
//                                           + (a%TM)*TM + b%TM],  s over (chunk, slice)
// grid = (n rows a, nsets), block 256.
__global__ void k_reduce_k(const double *__restrict__ kpart_all, int n, int tm, int ntiles,
                           const SyrkSets sets, int nchunk, double *__restrict__ klt,
                           int64_t ldk) {
  const int a = blockIdx.x, d = blockIdx.y;
  const int A = a / tm, al = a % tm;
  const int nsplit = sets.nsplit[d] * nchunk;  // slots (chunk, slice), chunk-major
  const double *__restrict__ kpart = kpart_all + sets.kbase[d];
  double *__restrict__ out = klt + d * ldk + tri0(a);
  for (int b = threadIdx.x; b <= a; b += blockDim.x) {
    if (nsplit == 0) {
      out[b] = 0.0;
      continue;
    }
    const int B = b / tm, bl = b % tm;
    const int64_t t = int64_t(A) * (A + 1) / 2 + B;
    const double *__restrict__ src = kpart + t * tm * tm + al * tm + bl;
    const int64_t stride = int64_t(ntiles) * tm * tm;
    double s = 0.0;
    int k = 0;
    for (; k + 8 <= nsplit; k += 8) {  // eight loads in flight, summed in slot order
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = src[(k + u) * stride];
#pragma unroll
      for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; k < nsplit; ++k) s += src[k * stride];
    out[b] = s;
  }
}

template <class P, int OT, int TA, int BK>
void launch_ht(Context *c, const typename P::T *L, const typename P::T *cb, int64_t j0, int64_t jc,
               typename P::T *X) {
  // two vectors per workgroup while the accumulators stay small (measured
  // 5-6% faster at OT = 4: half the orbital-fragment LDS reads per MFMA), and
  // at OT = 7 (4% faster; at OT = 5 and 6 it measured 3-7% slower: those keep
  // one vector, 6 in the 2 x 2 wave layout)
  constexpr int NJ = (OT * (TA / 64) <= 4 || (TA == 64 && OT == 7)) ? 2 : 1;
  const int ntile_a = c->n_pad / TA;
  const int64_t nblocks = ceil_div(jc, NJ) * ntile_a;
  const int64_t grid = round_up(nblocks, 8);
  if constexpr (std::is_same<P, F64>::value) {
    if (c->fv.p) {  // fused Coulomb pass (never together with screening)
      c->fv.ntile = ntile_a;
      hipLaunchKernelGGL((k_ht<P, OT, TA, BK, NJ, true>), dim3(unsigned(grid)), dim3(kThreads), 0,
                         c->stream, L, c->ldl, c->n, c->n_pad, cb, j0, jc, ntile_a, nblocks, X,
                         Screen{}, c->fv);
      return;
    }
  }
  hipLaunchKernelGGL((k_ht<P, OT, TA, BK, NJ, false>), dim3(unsigned(grid)), dim3(kThreads), 0,
                     c->stream, L, c->ldl, c->n, c->n_pad, cb, j0, jc, ntile_a, nblocks, X, c->scr,
                     FuseV{});
  if (c->scr.lnorm) c->x_tiles += jc * ntile_a;
}

// AO tile width: FP64 defaults to 64 (more workgroups per CU); FP32 uses 128
// while the accumulators fit (ot <= 6).  The context's ht_ta overrides
// (autotuner).
template <class P>
inline int ht_tile(const Context *c, int ot) {
  if (c->ht_ta == 64) return 64;
  if (c->ht_ta == 128 && (ot <= 6 || std::is_same<P, F32>::value)) return 128;
  if (std::is_same<P, F64>::value) return 64;
  return ot <= 6 ? 128 : 64;
}

// Wide banks: 128-wide AO tiles exist for FP32 only (the FP64 accumulators
// of 8+ tiles at 128 columns would not fit the register file).
template <class P, int OT>
void launch_ht_w(Context *c, bool wide, const typename P::T *L, const typename P::T *cb,
                 int64_t j0, int64_t jn, typename P::T *X) {
  // 8-row chunks where the staging pieces divide, else 16 (odd OT)
  constexpr int BK = (8 * 16 * OT) % kThreads == 0 ? 8 : 16;
  if constexpr (std::is_same<P, F32>::value) {
    if (wide) return launch_ht<P, OT, 128, BK>(c, L, cb, j0, jn, X);
  }
  launch_ht<P, OT, 64, BK>(c, L, cb, j0, jn, X);
}

template <class P>
void ht_dispatch(Context *c, int ot, const typename P::T *L, const typename P::T *cb, int64_t j0,
                 int64_t jn, typename P::T *X) {
  // BK = 8 where the staging pieces divide (16 for one or three tiles): small
  // LDS buffers let more workgroups share a CU, hiding the load latency behind
  // other workgroups' MFMAs (8% faster than BK = 16 at four tiles)
  const bool wide = ht_tile<P>(c, ot) == 128;
  switch (ot) {
    case 1: wide ? launch_ht<P, 1, 128, 16>(c, L, cb, j0, jn, X) : launch_ht<P, 1, 64, 16>(c, L, cb, j0, jn, X); break;
    case 2: wide ? launch_ht<P, 2, 128, 16>(c, L, cb, j0, jn, X) : launch_ht<P, 2, 64, 8>(c, L, cb, j0, jn, X); break;
    case 3: wide ? launch_ht<P, 3, 128, 16>(c, L, cb, j0, jn, X) : launch_ht<P, 3, 64, 16>(c, L, cb, j0, jn, X); break;
    case 4: wide ? launch_ht<P, 4, 128, 16>(c, L, cb, j0, jn, X) : launch_ht<P, 4, 64, 8>(c, L, cb, j0, jn, X); break;
    case 5: wide ? launch_ht<P, 5, 128, 16>(c, L, cb, j0, jn, X) : launch_ht<P, 5, 64, 16>(c, L, cb, j0, jn, X); break;
    case 6: wide ? launch_ht<P, 6, 128, 16>(c, L, cb, j0, jn, X) : launch_ht<P, 6, 64, 8>(c, L, cb, j0, jn, X); break;
    case 7: launch_ht_w<P, 7>(c, wide, L, cb, j0, jn, X); break;
    case 8: launch_ht_w<P, 8>(c, wide, L, cb, j0, jn, X); break;
    case 10: launch_ht_w<P, 10>(c, wide, L, cb, j0, jn, X); break;
    case 12: launch_ht_w<P, 12>(c, wide, L, cb, j0, jn, X); break;
    case 14: launch_ht_w<P, 14>(c, wide, L, cb, j0, jn, X); break;
    default: launch_ht_w<P, 16>(c, wide, L, cb, j0, jn, X); break;
  }
}

template <class P>
void syrk_launch(Context *c, int tm, int items, const typename P::T *X, int ldc, int64_t jn,
                 int ntiles, const SyrkSets &sets, double *kpart, int chunk) {
  if (tm == 128)
    hipLaunchKernelGGL((k_syrk<P, 128, 16>), dim3(unsigned(items)), dim3(kThreads), 0, c->stream, X,
                       c->n_pad, ldc, jn, ntiles, sets, kpart, chunk, c->scr);
  else
    hipLaunchKernelGGL((k_syrk<P, 64, 32>), dim3(unsigned(items)), dim3(kThreads), 0, c->stream, X,
                       c->n_pad, ldc, jn, ntiles, sets, kpart, chunk, c->scr);
  if (c->scr.xn2) c->k_items += items;
}

template <int OT, int TA>
void launch_ht_b3(Context *c, const float *L, const __bf16 *ch, const __bf16 *cl, int64_t j0,
                  int64_t jc, float *X) {
  const int ntile_a = c->n_pad / TA;
  const int64_t nblocks = jc * ntile_a;
  hipLaunchKernelGGL((k_ht_b3<OT, TA>), dim3(unsigned(round_up(nblocks, 8))), dim3(kThreads), 0,
                     c->stream, L, c->ldl, c->n, c->n_pad, ch, cl, j0, ntile_a, nblocks, X);
}

void ht_dispatch_b3(Context *c, int ot, const float *L, const __bf16 *ch, const __bf16 *cl,
                    int64_t j0, int64_t jn, float *X) {
  // 128-wide AO tiles unless the context's ht_ta asks for 64: twice the MFMA
  // work per staged chunk (measured 36-38% faster at 10 and 14 bank tiles)
  const bool wide = c->ht_ta != 64;
  switch (ot) {
    case 1: wide ? launch_ht_b3<1, 128>(c, L, ch, cl, j0, jn, X) : launch_ht_b3<1, 64>(c, L, ch, cl, j0, jn, X); break;
    case 2: wide ? launch_ht_b3<2, 128>(c, L, ch, cl, j0, jn, X) : launch_ht_b3<2, 64>(c, L, ch, cl, j0, jn, X); break;
    case 3: wide ? launch_ht_b3<3, 128>(c, L, ch, cl, j0, jn, X) : launch_ht_b3<3, 64>(c, L, ch, cl, j0, jn, X); break;
    case 4: wide ? launch_ht_b3<4, 128>(c, L, ch, cl, j0, jn, X) : launch_ht_b3<4, 64>(c, L, ch, cl, j0, jn, X); break;
    case 5: wide ? launch_ht_b3<5, 128>(c, L, ch, cl, j0, jn, X) : launch_ht_b3<5, 64>(c, L, ch, cl, j0, jn, X); break;
    case 6: wide ? launch_ht_b3<6, 128>(c, L, ch, cl, j0, jn, X) : launch_ht_b3<6, 64>(c, L, ch, cl, j0, jn, X); break;
    case 7: wide ? launch_ht_b3<7, 128>(c, L, ch, cl, j0, jn, X) : launch_ht_b3<7, 64>(c, L, ch, cl, j0, jn, X); break;
    case 8: wide ? launch_ht_b3<8, 128>(c, L, ch, cl, j0, jn, X) : launch_ht_b3<8, 64>(c, L, ch, cl, j0, jn, X); break;
    case 10: wide ? launch_ht_b3<10, 128>(c, L, ch, cl, j0, jn, X) : launch_ht_b3<10, 64>(c, L, ch, cl, j0, jn, X); break;
    case 12: wide ? launch_ht_b3<12, 128>(c, L, ch, cl, j0, jn, X) : launch_ht_b3<12, 64>(c, L, ch, cl, j0, jn, X); break;
    case 14: wide ? launch_ht_b3<14, 128>(c, L, ch, cl, j0, jn, X) : launch_ht_b3<14, 64>(c, L, ch, cl, j0, jn, X); break;
    default: wide ? launch_ht_b3<16, 128>(c, L, ch, cl, j0, jn, X) : launch_ht_b3<16, 64>(c, L, ch, cl, j0, jn, X); break;
  }
}

void syrk_launch_b3(Context *c, int tm, int items, const float *X, int ldc, int64_t jn, int ntiles,
                    const SyrkSets &sets, double *kpart, int chunk) {
  if (tm == 128)
    hipLaunchKernelGGL((k_syrk_b3<128>), dim3(unsigned(items)), dim3(kThreads), 0, c->stream, X,
                       c->n_pad, ldc, jn, ntiles, sets, kpart, chunk);
  else
    hipLaunchKernelGGL((k_syrk_b3<64>), dim3(unsigned(items)), dim3(kThreads), 0, c->stream, X,
                       c->n_pad, ldc, jn, ntiles, sets, kpart, chunk);
}

int pack_bank_b3(Context *c, int32_t nk, const int32_t *nocc, const double *const *corb,
                 const std::vector<int> &off, int ldc, DevBuf &buf) {
  const int n = c->n, n_pad = c->n_pad;
  const size_t half = sizeof(__bf16) * size_t(ldc) * n_pad;
  int rc;
  if ((rc = ensure(c, buf, 2 * half))) return rc;
  __bf16 *ch = static_cast<__bf16 *>(buf.ptr);
  __bf16 *cl = ch + size_t(ldc) * n_pad;
  phase_begin(c);
  if ((rc = fill_zero(c, ch, 2 * half))) return rc;
  for (int d = 0; d < nk; ++d) {
    if (nocc[d] == 0) continue;
    const int64_t cnt = int64_t(nocc[d]) * n;
    hipLaunchKernelGGL(k_pack_orb_b3, dim3(unsigned(ceil_div(cnt, 256))), dim3(256), 0, c->stream,
                       corb[d], nocc[d], n, n_pad, off[size_t(d)], ch, cl);
  }
  phase_end(c, MJK_PHASE_PACK);
  return MJK_OK;
}

// Layout of the bank for nk sets; returns MJK_OK and ot = 0 when empty.
int bank_layout(Context *c, int32_t nk, const int32_t *nocc, std::vector<int> &off,
                std::vector<int> &opad, int *ot_out, int *tot_out) {
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
  static const int kOT[] = {1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 14, 16};
  for (int v : kOT)
    if (16 * v >= tot) {
      *ot_out = v;
      return MJK_OK;
    }
  return set_error(c, MJK_ERR_ARG, "more than 256 orbital columns in one call");
}

template <typename T>
int pack_bank_t(Context *c, int32_t nk, const int32_t *nocc, const double *const *corb,
                const std::vector<int> &off, int ldc, DevBuf &buf) {
  const int n = c->n, n_pad = c->n_pad;
  const size_t bytes = sizeof(T) * size_t(n_pad) * ldc;
  int rc;
  if ((rc = ensure(c, buf, bytes))) return rc;
  T *cb = static_cast<T *>(buf.ptr);
  phase_begin(c);
  if ((rc = fill_zero(c, cb, bytes))) return rc;
  for (int d = 0; d < nk; ++d) {
    if (nocc[d] == 0) continue;
    const int64_t cnt = int64_t(nocc[d]) * n;
    hipLaunchKernelGGL(k_pack_orb<T>, dim3(unsigned(ceil_div(cnt, 256))), dim3(256), 0, c->stream,
                       corb[d], nocc[d], n, ldc, off[size_t(d)], cb);
  }
  phase_end(c, MJK_PHASE_PACK);
  return MJK_OK;
}

// FP32 copy of L, made once after the vectors change.
int ensure_l32(Context *c) {
  if (c->l32_valid) return MJK_OK;
  const int64_t cnt = c->nchol * c->ldl;
  int rc;
  if ((rc = ensure(c, c->ws_l32, sizeof(float) * size_t(cnt)))) return rc;
  phase_begin(c);
  hipLaunchKernelGGL(k_to_f32, dim3(4096), dim3(256), 0, c->stream, c->L,
                     static_cast<float *>(c->ws_l32.ptr), cnt);
  phase_end(c, MJK_PHASE_PACK);
  c->l32_valid = true;
  return hip_check(c, hipGetLastError(), "L to FP32");
}

// A run of vectors [j0, j1) computed in one precision.
struct Segment {
  int64_t j0, j1;
  int32_t prec;
};

int exchange_segments(Context *c, int32_t nk, const int32_t *nocc, const double *const *corb,
                      double *klt, int64_t ldk, const std::vector<Segment> &segs,
                      double tau = 0.0, double thr_x = 0.0) {
  const int n = c->n, n_pad = c->n_pad;
  int rc;
  if (nk > kMaxSets) return set_error(c, MJK_ERR_ARG, "exchange: more than 16 orbital sets");
  std::vector<int> off, opad;
  int ot = 0, tot = 0;
  if ((rc = bank_layout(c, nk, nocc, off, opad, &ot, &tot))) return rc;
  if (tot == 0) {  // no orbitals: K = 0
    for (int d = 0; d < nk; ++d)
      if ((rc = fill_zero(c, klt + d * ldk, sizeof(double) * c->np))) return rc;
    return MJK_OK;
  }
  const int ldc = 16 * ot;
  bool use[3] = {false, false, false};
  for (const Segment &s : segs)
    if (s.j1 > s.j0) use[s.prec] = true;
  const bool zsave = c->zs.z != nullptr;
  if (zsave && (nk != 1 || nocc[0] != c->zs.nocc))
    return set_error(c, MJK_ERR_ARG, "exchange: Z output needs exactly one orbital set");
  if (zsave && (tau > 0.0 || thr_x > 0.0))
    return set_error(c, MJK_ERR_ARG, "exchange: Z output cannot be screened");
  // one bank per precision in use; the reduced-precision ones read the FP32
  // copy of L.  Z = X C always multiplies by the FP64 bank.
  if ((use[MJK_PREC_FP64] || zsave) &&
      (rc = pack_bank_t<double>(c, nk, nocc, corb, off, ldc, c->w->cbank)))
    return rc;
  if (use[MJK_PREC_FP32] && (rc = pack_bank_t<float>(c, nk, nocc, corb, off, ldc, c->w->cb32)))
    return rc;
  if (use[MJK_PREC_BF16X3] && (rc = pack_bank_b3(c, nk, nocc, corb, off, ldc, c->w->cb3)))
    return rc;
  if ((use[MJK_PREC_FP32] || use[MJK_PREC_BF16X3]) && (rc = ensure_l32(c))) return rc;
  const float *L32 = static_cast<const float *>(c->ws_l32.ptr);

  // ---- chunking over vectors (X budget, sized for the widest element in use)
  const int64_t elem = use[MJK_PREC_FP64] ? 8 : 4;
  const int64_t xvec_bytes = elem * ldc * n_pad;
  int64_t jc = std::max<int64_t>(1, std::min<int64_t>(c->nchol, c->x_budget / xvec_bytes));
  int64_t nchunk = 0;
  for (const Segment &s : segs)
    if (s.j1 > s.j0) nchunk += ceil_div(s.j1 - s.j0, jc);
  if (jc * ldc >= (int64_t(1) << 31))
    return set_error(c, MJK_ERR_ARG, "exchange: chunk contraction length exceeds 2^31");
  if ((rc = ensure(c, c->w->x, size_t(xvec_bytes) * jc))) return rc;

  // ---- SYRK tiling and slices (one launch per chunk covers every density)
  int tm = c->tile_syrk ? c->tile_syrk : (n_pad >= 512 && n_pad % 128 == 0 ? 128 : 64);
  if (n_pad % tm) tm = 64;
  const int ntile1 = n_pad / tm;
  const int ntiles = ntile1 * (ntile1 + 1) / 2;
  SyrkSets sets{};
  sets.n = nk;
  int64_t kb = 0;
  sets.item0[0] = 0;
  // slices per density in proportion to its contraction length (opad rows per vector)
  int64_t work = 0;
  for (int d = 0; d < nk; ++d)
    if (nocc[d] > 0) work += opad[size_t(d)];
  // Workgroups per launch: at most the target, rounded down, so that a target of
  // whole resident rounds (k_syrk keeps two workgroups per CU) leaves no handful
  // of items to a round of its own.  Default: the most rounds (1, 2, 4 or 8) that
  // still give every item >= 2048 contraction rows; shorter items lose more to
  // their pipeline fill and store-out than finer balancing gains.
  const int64_t slots = 2 * int64_t(c->ncu);
  int rounds = 1;
  while (rounds < 8 && jc * work * ntiles >= int64_t(2048) * slots * (2 * rounds)) rounds *= 2;
  const int target = c->split_target ? c->split_target : int(slots * rounds);
  for (int d = 0; d < nk; ++d) {
    const int64_t want = int64_t(target) * opad[size_t(d)] /
                         (int64_t(ntiles) * std::max<int64_t>(work, 1));
    sets.nsplit[d] = nocc[d] > 0 ? int(std::max<int64_t>(1, std::min<int64_t>(want, jc))) : 0;
    sets.off[d] = off[size_t(d)];
    sets.opad[d] = std::max(opad[size_t(d)], 1);
    sets.kbase[d] = kb;
    kb += int64_t(sets.nsplit[d]) * nchunk * ntiles * tm * tm;
    sets.item0[d + 1] = sets.item0[d] + sets.nsplit[d] * ntiles;
  }
  const int items = sets.item0[nk];
  if ((rc = ensure(c, c->w->kpart, sizeof(double) * size_t(std::max<int64_t>(kb, 1))))) return rc;
  double *kpart = static_cast<double *>(c->w->kpart.ptr);

  // ---- screening set-up (norms of L strips, of the orbital sets, counters)
  const bool screen = tau > 0.0 || thr_x > 0.0;
  const int nstrip = n_pad / 64;
  c->k_items = c->x_tiles = 0;
  c->scr = Screen{};
  // the screening state must not outlive this call on any return path: other
  // users of the half-transform (puvx, backward) launch unscreened
  struct ScreenReset {
    Context *c;
    ~ScreenReset() { c->scr = Screen{}; }
  } screen_reset{c};
  if (screen) {
    if (use[MJK_PREC_BF16X3])
      return set_error(c, MJK_ERR_ARG, "exchange: screening is not available for split-BF16");
    if ((rc = ensure(c, c->scount, 2 * sizeof(unsigned long long))) ||
        (rc = fill_zero(c, c->scount.ptr, 2 * sizeof(unsigned long long))))
      return rc;
    c->scr.cnt = static_cast<unsigned long long *>(c->scount.ptr);
    c->scr.nstrip = nstrip;
    if (thr_x > 0.0) {
      if (!c->lnorm_valid) {
        if ((rc = ensure(c, c->lnorm, sizeof(double) * size_t(c->nchol) * nstrip))) return rc;
        hipLaunchKernelGGL(k_strip_norms, dim3(unsigned(c->nchol), unsigned(nstrip)), dim3(kThreads),
                           0, c->stream, c->L, c->ldl, n, nstrip,
                           static_cast<double *>(c->lnorm.ptr));
        c->lnorm_valid = true;
      }
      if ((rc = ensure(c, c->w->cnorm, sizeof(double) * (kMaxSets + 1)))) return rc;
      double *cn = static_cast<double *>(c->w->cnorm.ptr);
      if (use[MJK_PREC_FP64])
        hipLaunchKernelGGL(k_cnorm<double>, dim3(1), dim3(kThreads), 0, c->stream,
                           static_cast<const double *>(c->w->cbank.ptr), ldc, n, sets, cn);
      else
        hipLaunchKernelGGL(k_cnorm<float>, dim3(1), dim3(kThreads), 0, c->stream,
                           static_cast<const float *>(c->w->cb32.ptr), ldc, n, sets, cn);
      c->scr.lnorm = static_cast<const double *>(c->lnorm.ptr);
      c->scr.cmax = cn + nk;
      c->scr.thr_x = thr_x;
    }
    if (tau > 0.0) {
      if ((rc = ensure(c, c->w->xn2, sizeof(double) * size_t(nk) * jc * nstrip))) return rc;
      c->scr.tau = tau;
    }
  }

  int64_t slot = 0;  // running chunk index over all segments
  for (const Segment &s : segs) {
    for (int64_t j0 = s.j0; j0 < s.j1; j0 += jc, ++slot) {
      const int64_t jn = std::min(jc, s.j1 - j0);
      phase_begin(c);
      switch (s.prec) {
        case MJK_PREC_FP64:
          ht_dispatch<F64>(c, ot, c->L, static_cast<const double *>(c->w->cbank.ptr), j0, jn,
                           static_cast<double *>(c->w->x.ptr));
          break;
        case MJK_PREC_FP32:
          ht_dispatch<F32>(c, ot, L32, static_cast<const float *>(c->w->cb32.ptr), j0, jn,
                           static_cast<float *>(c->w->x.ptr));
          break;
        default: {
          const __bf16 *cbh = static_cast<const __bf16 *>(c->w->cb3.ptr);
          ht_dispatch_b3(c, ot, L32, cbh, cbh + size_t(ldc) * n_pad, j0, jn,
                         static_cast<float *>(c->w->x.ptr));
        }
      }
      phase_end(c, MJK_PHASE_HALF_TRANSFORM);
      if (zsave) {
        phase_begin(c);
        if ((rc = z_from_x(c, c->w->x.ptr, s.prec != MJK_PREC_FP64, ldc, nocc[0], jn,
                           c->zs.z + j0 * int64_t(nocc[0]) * nocc[0])))
          return rc;
        phase_end(c, MJK_PHASE_ZSAVE);
      }
      if (screen && tau > 0.0) {
        double *xn2 = static_cast<double *>(c->w->xn2.ptr);
        if (s.prec == MJK_PREC_FP64)
          hipLaunchKernelGGL(k_xnorms<double>, dim3(unsigned(jn), unsigned(nstrip), unsigned(nk)),
                             dim3(kThreads), 0, c->stream, static_cast<const double *>(c->w->x.ptr),
                             ldc, n_pad, sets, nstrip, xn2);
        else
          hipLaunchKernelGGL(k_xnorms<float>, dim3(unsigned(jn), unsigned(nstrip), unsigned(nk)),
                             dim3(kThreads), 0, c->stream, static_cast<const float *>(c->w->x.ptr),
                             ldc, n_pad, sets, nstrip, xn2);
        c->scr.xn2 = xn2;
      }
      phase_begin(c);
      switch (s.prec) {
        case MJK_PREC_FP64:
          syrk_launch<F64>(c, tm, items, static_cast<const double *>(c->w->x.ptr), ldc, jn, ntiles,
                           sets, kpart, int(slot));
          break;
        case MJK_PREC_FP32:
          syrk_launch<F32>(c, tm, items, static_cast<const float *>(c->w->x.ptr), ldc, jn, ntiles,
                           sets, kpart, int(slot));
          break;
        default:
          syrk_launch_b3(c, tm, items, static_cast<const float *>(c->w->x.ptr), ldc, jn, ntiles,
                         sets, kpart, int(slot));
      }
      phase_end(c, MJK_PHASE_EXCHANGE);
    }
  }
  phase_begin(c);
  hipLaunchKernelGGL(k_reduce_k, dim3(unsigned(n), unsigned(nk)), dim3(256), 0, c->stream, kpart, n,
                     tm, ntiles, sets, int(nchunk), klt, ldk);
  phase_end(c, MJK_PHASE_REDUCE);
  return hip_check(c, hipGetLastError(), "exchange launch");
}

}  // namespace

int pack_bank(Context *c, int32_t nk, const int32_t *nocc, const double *const *corb,
              std::vector<int> &off, std::vector<int> &opad, int *ot, int *tot) {
  int rc = bank_layout(c, nk, nocc, off, opad, ot, tot);
  if (rc || *tot == 0) return rc;
  return pack_bank_t<double>(c, nk, nocc, corb, off, 16 * *ot, c->w->cbank);
}

void ht_launch(Context *c, int ot, const double *cb, int64_t j0, int64_t jn, double *X) {
  ht_dispatch<F64>(c, ot, c->L, cb, j0, jn, X);
}

int exchange(Context *c, int32_t nk, const int32_t *nocc, const double *const *corb, double *klt,
             int64_t ldk, int32_t precision) {
  const int64_t M = c->nchol;
  if (precision == MJK_PREC_LADDER)  // split points of mjk_set_ladder; FP64 until set
    return exchange_mixed(c, nk, nocc, corb, klt, ldk, c->lad_fp32 < 0 ? M : c->lad_fp32,
                          c->lad_bf16 < 0 ? M : c->lad_bf16);
  if (precision < MJK_PREC_FP64 || precision > MJK_PREC_BF16X3)
    return set_error(c, MJK_ERR_ARG, "exchange: precision mode not available");
  return exchange_mixed(c, nk, nocc, corb, klt, ldk, precision == MJK_PREC_FP64 ? M : 0,
                        precision == MJK_PREC_BF16X3 ? 0 : M);
}

int exchange_screened(Context *c, int32_t nk, const int32_t *nocc, const double *const *corb,
                      double *klt, int64_t ldk, int32_t precision, double tau, double thr_x) {
  if (nk < 1) return set_error(c, MJK_ERR_ARG, "exchange: nk < 1");
  if (ldk < c->np) return set_error(c, MJK_ERR_ARG, "exchange: ldk < nbas(nbas+1)/2");
  if (c->nloaded != c->nchol)
    return set_error(c, MJK_ERR_SHORT_READ, "exchange: not all Cholesky vectors were uploaded");
  if (!(tau >= 0.0) || !(thr_x >= 0.0))
    return set_error(c, MJK_ERR_ARG, "exchange: screening thresholds must be >= 0");
  if (precision == MJK_PREC_LADDER) {
    if (tau > 0.0 || thr_x > 0.0)
      return set_error(c, MJK_ERR_ARG, "exchange: the precision ladder cannot be screened");
    return exchange(c, nk, nocc, corb, klt, ldk, precision);
  }
  if (precision < MJK_PREC_FP64 || precision > MJK_PREC_BF16X3)
    return set_error(c, MJK_ERR_ARG, "exchange: precision mode not available");
  std::vector<Segment> segs = {{0, c->nchol, precision}};
  return exchange_segments(c, nk, nocc, corb, klt, ldk, segs, tau, thr_x);
}

int exchange_mixed(Context *c, int32_t nk, const int32_t *nocc, const double *const *corb,
                   double *klt, int64_t ldk, int64_t j_fp32, int64_t j_bf16) {
  if (nk < 1) return set_error(c, MJK_ERR_ARG, "exchange: nk < 1");
  if (ldk < c->np) return set_error(c, MJK_ERR_ARG, "exchange: ldk < nbas(nbas+1)/2");
  if (c->nloaded != c->nchol)
    return set_error(c, MJK_ERR_SHORT_READ, "exchange: not all Cholesky vectors were uploaded");
  const int64_t M = c->nchol;
  if (j_fp32 < 0 || j_fp32 > j_bf16 || j_bf16 > M)
    return set_error(c, MJK_ERR_ARG, "exchange: need 0 <= j_fp32 <= j_bf16 <= nchol");
  std::vector<Segment> segs = {{0, j_fp32, MJK_PREC_FP64},
                               {j_fp32, j_bf16, MJK_PREC_FP32},
                               {j_bf16, M, MJK_PREC_BF16X3}};
  return exchange_segments(c, nk, nocc, corb, klt, ldk, segs);
}

}  // namespace mjk
