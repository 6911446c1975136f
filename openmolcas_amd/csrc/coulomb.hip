// Coulomb (J) build from HBM-resident packed Cholesky vectors.
//
// Reference: the DGEMV pair of src/fock_util/cho_fscf.F90:208,215 (also
// cho_lk_scf.F90:442-449, cho_fmcscf.F90:261-285):
//   V_J  = sum_p L^J_p D_p          (D folded: off-diagonal doubled)
//   J_p  = sum_J L^J_p V_J
// Both passes stream L once; they are HBM-bound.  L is stored vector-major
// (row J holds the packed LT of L^J, stride ldl, a multiple of 512 doubles),
// so each pass reads whole 4 KiB row segments per workgroup iteration.
#include "mjk_internal.h"

namespace mjk {
namespace {

constexpr int kThreads = 256;
constexpr int kSeg = 2 * kThreads;  // doubles per row segment per iteration

template <int N>
__device__ inline void block_sum(double (&v)[N], double *lds /* [N][4] */) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    double x = v[i];
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) x += __shfl_down(x, off, kWave);
    if (lane == 0) lds[i * 4 + wave] = x;
  }
  __syncthreads();
  if (threadIdx.x < N) {
    const double *p = lds + threadIdx.x * 4;
    v[0] = (p[0] + p[1]) + (p[2] + p[3]);
  }
}

// Pass 1.  grid = (ceil(M/JB), nsplit).  Workgroup (jb, s) accumulates
// V(J, d) for JB vectors over the pair range of split s into
// vpart[(s*M + J)*nd + d].
template <int JB, int ND>
__global__ __launch_bounds__(kThreads) void k_coulomb_v(
    const double *__restrict__ L, int64_t ldl, int64_t M, const double *__restrict__ D,
    int nd, int64_t pchunk, double *__restrict__ vpart) {
  __shared__ double lds[JB * ND * 4];
  const int64_t j0 = int64_t(blockIdx.x) * JB;
  const int64_t p0 = int64_t(blockIdx.y) * pchunk;
  const int64_t p1 = min(p0 + pchunk, ldl);
  double acc[JB * ND];
#pragma unroll
  for (int i = 0; i < JB * ND; ++i) acc[i] = 0.0;
  for (int64_t p = p0 + 2 * threadIdx.x; p < p1; p += kSeg) {
    v2d dv[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d)
      dv[d] = (d < nd) ? *reinterpret_cast<const v2d *>(D + d * ldl + p)
                       : v2d{0.0, 0.0};
#pragma unroll
    for (int jb = 0; jb < JB; ++jb) {
      if (j0 + jb < M) {
        const v2d l = __builtin_nontemporal_load(
            reinterpret_cast<const v2d *>(L + (j0 + jb) * ldl + p));
#pragma unroll
        for (int d = 0; d < ND; ++d) acc[jb * ND + d] += l.x * dv[d].x + l.y * dv[d].y;
      }
    }
  }
  block_sum<JB * ND>(acc, lds);
  if (threadIdx.x < JB * ND) {
    const int jb = threadIdx.x / ND, d = threadIdx.x % ND;
    if (j0 + jb < M && d < nd) vpart[(int64_t(blockIdx.y) * M + j0 + jb) * nd + d] = acc[0];
  }
}

// V(J, d) = sum_s vpart[(s*M + J)*nd + d]  (fixed order: deterministic).
__global__ void k_reduce_v(const double *__restrict__ vpart, int64_t M, int nd, int nsplit,
                           double *__restrict__ v) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= M * nd) return;
  double s = 0.0;
  for (int k = 0; k < nsplit; ++k) s += vpart[int64_t(k) * M * nd + i];
  v[i] = s;
}

// Pass 2.  grid = (ceil(ldl/(W*kSeg)), nsplit).  Workgroup (c, s) owns pairs
// [c*W*kSeg, (c+1)*W*kSeg) (W row segments, so each vector row is read in
// W*4 KiB runs) and sums over the vectors of split s:
// jpart[(s*nd + d)*ldl + p] = sum_J L(J,p) V(J,d).  U rows in flight per thread.
template <int ND, int W, int U>
__global__ __launch_bounds__(kThreads) void k_coulomb_j(
    const double *__restrict__ L, int64_t ldl, int64_t M, const double *__restrict__ V,
    int nd, int64_t jchunk, double *__restrict__ jpart) {
  const int64_t pb = int64_t(blockIdx.x) * (W * kSeg) + 2 * threadIdx.x;
  const int64_t ja = int64_t(blockIdx.y) * jchunk;
  const int64_t jz = min(ja + jchunk, M);
  v2d acc[W][ND];
  bool ok[W];
#pragma unroll
  for (int w = 0; w < W; ++w) {
    ok[w] = pb + w * kSeg < ldl;
#pragma unroll
    for (int d = 0; d < ND; ++d) acc[w][d] = v2d{0.0, 0.0};
  }
  int64_t j = ja;
  for (; j + U <= jz; j += U) {
    v2d l[U][W];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int w = 0; w < W; ++w)
        l[u][w] = ok[w] ? __builtin_nontemporal_load(
                              reinterpret_cast<const v2d *>(L + (j + u) * ldl + pb + w * kSeg))
                        : v2d{0.0, 0.0};
#pragma unroll
    for (int u = 0; u < U; ++u) {
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        if (d < nd) {
          const double v = V[(j + u) * nd + d];
#pragma unroll
          for (int w = 0; w < W; ++w) {
            acc[w][d].x += l[u][w].x * v;
            acc[w][d].y += l[u][w].y * v;
          }
        }
      }
    }
  }
  for (; j < jz; ++j) {
#pragma unroll
    for (int w = 0; w < W; ++w) {
      if (!ok[w]) continue;
      const v2d l = __builtin_nontemporal_load(reinterpret_cast<const v2d *>(L + j * ldl + pb + w * kSeg));
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        if (d < nd) {
          const double v = V[j * nd + d];
          acc[w][d].x += l.x * v;
          acc[w][d].y += l.y * v;
        }
      }
    }
  }
#pragma unroll
  for (int w = 0; w < W; ++w)
#pragma unroll
    for (int d = 0; d < ND; ++d)
      if (ok[w] && d < nd)
        *reinterpret_cast<v2d *>(jpart + (int64_t(blockIdx.y) * nd + d) * ldl + pb + w * kSeg) =
            acc[w][d];
}

// jlt(d*ldj + p) = sum_s jpart[(s*nd + d)*ldl + p], p < np.
__global__ void k_reduce_j(const double *__restrict__ jpart, int64_t ldl, int64_t np, int nd,
                           int nsplit, double *__restrict__ jlt, int64_t ldj) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int d = blockIdx.y;
  if (p >= np) return;
  double s = 0.0;
  for (int k = 0; k < nsplit; ++k) s += jpart[(int64_t(k) * nd + d) * ldl + p];
  jlt[d * ldj + p] = s;
}

// Copies nd densities of stride ldd into the zero-padded staging buffer of
// stride ldl.
__global__ void k_pack_dens(const double *__restrict__ dlt, int64_t ldd, int64_t np,
                            int64_t ldl, double *__restrict__ out) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int d = blockIdx.y;
  if (p >= ldl) return;
  out[d * ldl + p] = (p < np) ? dlt[d * ldd + p] : 0.0;
}


// ---- many densities (3..16): both passes as FP64-MFMA GEMMs ---------------
// With nd densities each vector element is used nd times; the per-vector
// kernels above would re-read every density for every vector.  Here a
// workgroup stages a tile of L and a tile of the densities (or of V) in LDS and
// contracts them with v_mfma_f64_16x16x4f64 (A[l%16][l/16], B[l/16][l%16],
// result row 4*r + l/16, column l%16), so each density element is read once
// per 64 vectors and each L element once per pass: still HBM-bound, at the
// single-density byte count.
constexpr int kVM = 64;   // vectors per workgroup (pass 1)
constexpr int kVK = 32;   // pairs per LDS chunk (pass 1)
constexpr int kJP = 256;  // pairs per workgroup (pass 2)
constexpr int kJK = 16;   // vectors per LDS chunk (pass 2)

__device__ inline v4d mfma64(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// grid = (ceil(M/64), nsplit); vpart[(s*M + J)*nd + d] = sum_{p in split s} L[J][p] D[d][p]
__global__ __launch_bounds__(kThreads) void k_coulomb_v_mfma(
    const double *__restrict__ L, int64_t ldl, int64_t M, const double *__restrict__ D, int nd,
    int64_t pchunk, double *__restrict__ vpart) {
  __shared__ double sL[kVM][kVK + 1];
  __shared__ double sD[16][kVK + 1];
  const int64_t J0 = int64_t(blockIdx.x) * kVM;
  const int64_t p0 = int64_t(blockIdx.y) * pchunk;
  const int64_t p1 = min(p0 + pchunk, ldl);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  v4d acc = {0.0, 0.0, 0.0, 0.0};
  constexpr int NL = kVM * kVK / kThreads, NDD = 16 * kVK / kThreads;
  double rl[NL], rd[NDD];
  auto fetch = [&](int64_t pc) {
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      const int e = tid + kThreads * i, r = e / kVK, k = e % kVK;
      const bool ok = (J0 + r < M);
      const double v = L[(ok ? J0 + r : 0) * ldl + pc + k];
      rl[i] = ok ? v : 0.0;
    }
#pragma unroll
    for (int i = 0; i < NDD; ++i) {
      const int e = tid + kThreads * i, d = e / kVK, k = e % kVK;
      const double v = D[(d < nd ? d : 0) * ldl + pc + k];
      rd[i] = d < nd ? v : 0.0;
    }
  };
  if (p0 < p1) fetch(p0);
  for (int64_t pc = p0; pc < p1; pc += kVK) {
    if (pc != p0) __syncthreads();
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      const int e = tid + kThreads * i;
      sL[e / kVK][e % kVK] = rl[i];
    }
#pragma unroll
    for (int i = 0; i < NDD; ++i) {
      const int e = tid + kThreads * i;
      sD[e / kVK][e % kVK] = rd[i];
    }
    __syncthreads();
    if (pc + kVK < p1) fetch(pc + kVK);
#pragma unroll
    for (int ks = 0; ks < kVK / 4; ++ks) {
      const int k = ks * 4 + (lane >> 4);
      acc = mfma64(sL[16 * w + (lane & 15)][k], sD[lane & 15][k], acc);
    }
  }
  const int d = lane & 15;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int64_t J = J0 + 16 * w + 4 * r + (lane >> 4);
    if (J < M && d < nd) vpart[(int64_t(blockIdx.y) * M + J) * nd + d] = acc[r];
  }
}

// grid = (ldl/256, nsplit); jpart[(s*nd + d)*ldl + p] = sum_{J in split s} V[J][d] L[J][p]
__global__ __launch_bounds__(kThreads) void k_coulomb_j_mfma(
    const double *__restrict__ L, int64_t ldl, int64_t M, const double *__restrict__ V, int nd,
    int64_t jchunk, double *__restrict__ jpart) {
  __shared__ double sL[kJK][kJP];
  __shared__ double sV[kJK][16];
  const int64_t p0 = int64_t(blockIdx.x) * kJP;
  const int64_t ja = int64_t(blockIdx.y) * jchunk;
  const int64_t jz = min(ja + jchunk, M);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  v4d acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = v4d{0.0, 0.0, 0.0, 0.0};
  constexpr int NL = kJK * kJP / kThreads;
  double rl[NL], rv;
  auto fetch = [&](int64_t jc) {
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      const int e = tid + kThreads * i, r = e / kJP, k = e % kJP;
      const bool ok = jc + r < jz;
      const double v = L[(ok ? jc + r : 0) * ldl + p0 + k];
      rl[i] = ok ? v : 0.0;
    }
    const int r = tid / 16, d = tid % 16;  // 256 threads = 16 x 16
    const bool ok = (jc + r < jz) && d < nd;
    const double v = V[(ok ? (jc + r) * nd + d : 0)];
    rv = ok ? v : 0.0;
  };
  if (ja < jz) fetch(ja);
  for (int64_t jc = ja; jc < jz; jc += kJK) {
    if (jc != ja) __syncthreads();
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      const int e = tid + kThreads * i;
      sL[e / kJP][e % kJP] = rl[i];
    }
    sV[tid / 16][tid % 16] = rv;
    __syncthreads();
    if (jc + kJK < jz) fetch(jc + kJK);
#pragma unroll
    for (int ks = 0; ks < kJK / 4; ++ks) {
      const int k = ks * 4 + (lane >> 4);
      const double b = sV[k][lane & 15];
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[t] = mfma64(sL[k][64 * w + 16 * t + (lane & 15)], b, acc[t]);
    }
  }
  const int d = lane & 15;
  if (d < nd) {
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t p = p0 + 64 * w + 16 * t + 4 * r + (lane >> 4);
        jpart[(int64_t(blockIdx.y) * nd + d) * ldl + p] = acc[t][r];
      }
  }
}

template <int JB, int ND>
void launch_v(Context *c, const double *D, int nd, double *vpart, int nsplit, int64_t pchunk) {
  dim3 grid(unsigned(ceil_div(c->nchol, JB)), unsigned(nsplit));
  hipLaunchKernelGGL((k_coulomb_v<JB, ND>), grid, dim3(kThreads), 0, c->stream, c->L, c->ldl,
                     c->nchol, D, nd, pchunk, vpart);
}

// Row segments per workgroup of the J pass: two for one or two densities
// (8 KiB runs per vector row, measured faster), one otherwise (register budget).
constexpr int j_width(int ndt) { return ndt <= 2 ? 2 : 1; }

template <int ND>
void launch_j(Context *c, const double *V, int nd, double *jpart, int nsplit, int64_t jchunk) {
  constexpr int W = j_width(ND), U = 4;
  dim3 grid(unsigned(ceil_div(c->ldl / kSeg, W)), unsigned(nsplit));
  hipLaunchKernelGGL((k_coulomb_j<ND, W, U>), grid, dim3(kThreads), 0, c->stream, c->L, c->ldl,
                     c->nchol, V, nd, jchunk, jpart);
}

// w[J] = ||L^J||_F^2 of the full symmetric matrix = 2 sum_p L_p^2 - sum_a L_aa^2.
// grid = M, one workgroup per vector.
__global__ __launch_bounds__(kThreads) void k_vector_norms(const double *__restrict__ L,
                                                           int64_t ldl, int64_t np, int n,
                                                           double *__restrict__ w) {
  __shared__ double lds[4];
  const double *__restrict__ row = L + int64_t(blockIdx.x) * ldl;
  double s[1] = {0.0};
  for (int64_t p = threadIdx.x; p < np; p += kThreads) s[0] += 2.0 * row[p] * row[p];
  for (int a = threadIdx.x; a < n; a += kThreads) {
    const double x = row[int64_t(a) * (a + 1) / 2 + a];
    s[0] -= x * x;
  }
  block_sum<1>(s, lds);
  if (threadIdx.x == 0) w[blockIdx.x] = s[0];
}

}  // namespace

int vector_norms(Context *c, double *w) {
  if (!w) return set_error(c, MJK_ERR_ARG, "vector_norms: null pointer");
  if (c->nloaded != c->nchol)
    return set_error(c, MJK_ERR_SHORT_READ, "vector_norms: not all Cholesky vectors were uploaded");
  hipLaunchKernelGGL(k_vector_norms, dim3(unsigned(c->nchol)), dim3(kThreads), 0, c->stream, c->L,
                     c->ldl, c->np, c->n, w);
  return hip_check(c, hipGetLastError(), "vector_norms launch");
}

int coulomb(Context *c, int32_t nd, const double *dlt, int64_t ldd, double *jlt, int64_t ldj,
            double *v_out) {
  if (nd < 1 || nd > 16) return set_error(c, MJK_ERR_ARG, "coulomb: nd must be in [1,16]");
  if (ldd < c->np || ldj < c->np) return set_error(c, MJK_ERR_ARG, "coulomb: ld < nbas(nbas+1)/2");
  if (c->nloaded != c->nchol)
    return set_error(c, MJK_ERR_SHORT_READ, "coulomb: not all Cholesky vectors were uploaded");
  const int64_t M = c->nchol, ldl = c->ldl;
  int rc;

  // staging copy of the densities (zero padded to ldl)
  if ((rc = ensure(c, c->w->d, sizeof(double) * nd * ldl))) return rc;
  double *Dp = static_cast<double *>(c->w->d.ptr);
  phase_begin(c);
  hipLaunchKernelGGL(k_pack_dens, dim3(unsigned(ceil_div(ldl, 256)), unsigned(nd)), dim3(256), 0,
                     c->stream, dlt, ldd, c->np, ldl, Dp);
  phase_end(c, MJK_PHASE_PACK);

  const int64_t nseg = ldl / kSeg;
  if (nd >= 3) {
    // ---- pass 1 (MFMA): V = L D^T, split over pairs to fill the device
    const int64_t nblk = ceil_div(M, kVM);
    const int64_t nk = ldl / kVK;
    int nsplit1 = int(std::min<int64_t>(std::max<int64_t>(ceil_div(2048, nblk), 1), nk));
    const int64_t pchunk = ceil_div(nk, nsplit1) * kVK;
    nsplit1 = int(ceil_div(ldl, pchunk));
    if ((rc = ensure(c, c->w->vpart, sizeof(double) * nsplit1 * M * nd))) return rc;
    if ((rc = ensure(c, c->w->v, sizeof(double) * M * nd))) return rc;
    double *vpart = static_cast<double *>(c->w->vpart.ptr);
    double *V = v_out ? v_out : static_cast<double *>(c->w->v.ptr);
    phase_begin(c);
    hipLaunchKernelGGL(k_coulomb_v_mfma, dim3(unsigned(nblk), unsigned(nsplit1)), dim3(kThreads), 0,
                       c->stream, c->L, ldl, M, Dp, nd, pchunk, vpart);
    phase_end(c, MJK_PHASE_COULOMB_V);
    phase_begin(c);
    hipLaunchKernelGGL(k_reduce_v, dim3(unsigned(ceil_div(M * nd, 256))), dim3(256), 0, c->stream,
                       vpart, M, nd, nsplit1, V);
    phase_end(c, MJK_PHASE_REDUCE);
    // ---- pass 2 (MFMA): J = L^T V, split over vectors
    const int64_t npc = ldl / kJP;
    int nsplit2 = int(std::min<int64_t>(std::max<int64_t>(ceil_div(2048, npc), 1), ceil_div(M, kJK)));
    const int64_t jchunk = round_up(ceil_div(M, nsplit2), kJK);
    nsplit2 = int(ceil_div(M, jchunk));
    if ((rc = ensure(c, c->w->jpart, sizeof(double) * nsplit2 * nd * ldl))) return rc;
    double *jpart = static_cast<double *>(c->w->jpart.ptr);
    phase_begin(c);
    hipLaunchKernelGGL(k_coulomb_j_mfma, dim3(unsigned(npc), unsigned(nsplit2)), dim3(kThreads), 0,
                       c->stream, c->L, ldl, M, V, nd, jchunk, jpart);
    phase_end(c, MJK_PHASE_COULOMB_J);
    phase_begin(c);
    hipLaunchKernelGGL(k_reduce_j, dim3(unsigned(ceil_div(c->np, 256)), unsigned(nd)), dim3(256), 0,
                       c->stream, jpart, ldl, c->np, nd, nsplit2, jlt, ldj);
    phase_end(c, MJK_PHASE_REDUCE);
    return hip_check(c, hipGetLastError(), "coulomb launch");
  }

  // ---- pass 1: V = L^T D
  const int ndt = nd <= 1 ? 1 : nd <= 2 ? 2 : nd <= 4 ? 4 : nd <= 8 ? 8 : 16;
  const int jb = ndt <= 2 ? 8 : ndt == 4 ? 4 : ndt == 8 ? 2 : 1;
  const int64_t nblk = ceil_div(M, jb);
  int nsplit1 = int(std::min<int64_t>(std::max<int64_t>(ceil_div(2048, nblk), 1), nseg));
  const int64_t pchunk = ceil_div(nseg, nsplit1) * kSeg;
  nsplit1 = int(ceil_div(ldl, pchunk));
  if ((rc = ensure(c, c->w->vpart, sizeof(double) * nsplit1 * M * nd))) return rc;
  if ((rc = ensure(c, c->w->v, sizeof(double) * M * nd))) return rc;
  double *vpart = static_cast<double *>(c->w->vpart.ptr);
  double *V = v_out ? v_out : static_cast<double *>(c->w->v.ptr);
  phase_begin(c);
  switch (ndt) {
    case 1: launch_v<8, 1>(c, Dp, nd, vpart, nsplit1, pchunk); break;
    case 2: launch_v<8, 2>(c, Dp, nd, vpart, nsplit1, pchunk); break;
    case 4: launch_v<4, 4>(c, Dp, nd, vpart, nsplit1, pchunk); break;
    case 8: launch_v<2, 8>(c, Dp, nd, vpart, nsplit1, pchunk); break;
    default: launch_v<1, 16>(c, Dp, nd, vpart, nsplit1, pchunk); break;
  }
  phase_end(c, MJK_PHASE_COULOMB_V);
  phase_begin(c);
  hipLaunchKernelGGL(k_reduce_v, dim3(unsigned(ceil_div(M * nd, 256))), dim3(256), 0, c->stream,
                     vpart, M, nd, nsplit1, V);
  phase_end(c, MJK_PHASE_REDUCE);

  // ---- pass 2: J = L V (about 1024 workgroups of two segments, 2048 of one)
  const int64_t nxj = ceil_div(nseg, j_width(ndt));
  int nsplit2 = int(std::min<int64_t>(std::max<int64_t>(ceil_div(2048 / j_width(ndt), nxj), 1), M));
  const int64_t jchunk = ceil_div(M, nsplit2);
  nsplit2 = int(ceil_div(M, jchunk));
  if ((rc = ensure(c, c->w->jpart, sizeof(double) * nsplit2 * nd * ldl))) return rc;
  double *jpart = static_cast<double *>(c->w->jpart.ptr);
  phase_begin(c);
  switch (ndt) {
    case 1: launch_j<1>(c, V, nd, jpart, nsplit2, jchunk); break;
    case 2: launch_j<2>(c, V, nd, jpart, nsplit2, jchunk); break;
    case 4: launch_j<4>(c, V, nd, jpart, nsplit2, jchunk); break;
    case 8: launch_j<8>(c, V, nd, jpart, nsplit2, jchunk); break;
    default: launch_j<16>(c, V, nd, jpart, nsplit2, jchunk); break;
  }
  phase_end(c, MJK_PHASE_COULOMB_J);
  phase_begin(c);
  hipLaunchKernelGGL(k_reduce_j, dim3(unsigned(ceil_div(c->np, 256)), unsigned(nd)), dim3(256), 0,
                     c->stream, jpart, ldl, c->np, nd, nsplit2, jlt, ldj);
  phase_end(c, MJK_PHASE_REDUCE);
  return hip_check(c, hipGetLastError(), "coulomb launch");
}

}  // namespace mjk
