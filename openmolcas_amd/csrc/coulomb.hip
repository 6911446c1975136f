// Coulomb (J) build from HBM-resident packed Cholesky vectors.
//
// Reference: the DGEMV pair of src/fock_util/cho_fscf.F90:208,215 (also
// cho_lk_scf.F90:442-449, cho_fmcscf.F90:261-285):
//   V_J  = sum_p L^J_p D_p          (D folded: off-diagonal doubled)
//   J_p  = sum_J L^J_p V_J
// Both passes stream L once; they are HBM-bound.  L is stored vector-major
// (row J holds the packed LT of L^J, stride ldl, a multiple of 512 doubles),
// so each pass reads whole 4 KiB row segments per workgroup iteration.
#include <mutex>
#include <string>

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
    // rows past M read the last row (their sums are never written): the loads
    // stay unconditional, so all JB of them are in flight together
#pragma unroll
    for (int jb = 0; jb < JB; ++jb) {
      const int64_t j = j0 + jb < M ? j0 + jb : M - 1;
      const v2d l = __builtin_nontemporal_load(reinterpret_cast<const v2d *>(L + j * ldl + p));
#pragma unroll
      for (int d = 0; d < ND; ++d) acc[jb * ND + d] += l.x * dv[d].x + l.y * dv[d].y;
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

// ---- many densities (3..16): both passes on FP64 MFMA ---------------------
// With nd densities each vector element is used nd times; the per-vector
// kernels above would re-read every density for every vector.  Both passes
// below read each L element once per pass and each density (or V) element
// once per 64 vectors: still HBM-bound, at the single-density byte count.
// MFMA v_mfma_f64_16x16x4f64: lane l supplies A[l%16][l/16] and B[l/16][l%16];
// result register r holds row 4r + l/16, column l%16.
__device__ inline v4d mfma64(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Pass 1 as V = L D^T.  grid = (ceil(M/VM), nsplit); workgroup (x, s) stages
// VM vectors x VK pairs of L (rows read in VK*8-byte runs with 16-byte loads)
// and the same VK pairs of the densities in LDS, the next chunk being fetched
// to registers meanwhile.  Row tile rt = w / WPT of 16 vectors is contracted
// by WPT waves, each taking every WPT-th group of k-steps; their partial
// tiles are summed through LDS.  vpart[(s*M + J)*nd + d].
template <int VM, int VK>
__global__ __launch_bounds__(kThreads) void k_coulomb_v_tile(
    const double *__restrict__ L, int64_t ldl, int64_t M, const double *__restrict__ D, int nd,
    int64_t pchunk, double *__restrict__ vpart) {
  constexpr int RT = VM / 16, WPT = 4 / RT, KS = VK / 4 / WPT;
  constexpr int NL = VM * VK / 2 / kThreads, NDD = 16 * VK / 2 / kThreads;
  static_assert(NL >= 1 && NDD >= 1 && KS >= 1, "tile shape");
  __shared__ double sL[VM][VK + 1];
  __shared__ double sD[16][VK + 1];
  const int64_t J0 = int64_t(blockIdx.x) * VM;
  const int64_t p0 = int64_t(blockIdx.y) * pchunk;
  const int64_t p1 = min(p0 + pchunk, ldl);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int rt = w / WPT, wk = w % WPT;
  v4d acc = {0.0, 0.0, 0.0, 0.0};
  v2d rl[NL], rd[NDD];
  // Out-of-range rows (vectors >= M, densities >= nd) load row 0 and are
  // zeroed when stashed, so the chunk's loads stay in flight across the
  // multiply of the previous chunk.
  auto fetch = [&](int64_t pc) {
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      const int e = tid + kThreads * i, r = e / (VK / 2), k2 = e % (VK / 2);
      rl[i] = __builtin_nontemporal_load(
          reinterpret_cast<const v2d *>(L + (J0 + r < M ? J0 + r : 0) * ldl + pc + 2 * k2));
    }
#pragma unroll
    for (int i = 0; i < NDD; ++i) {
      const int e = tid + kThreads * i, d = e / (VK / 2), k2 = e % (VK / 2);
      rd[i] = *reinterpret_cast<const v2d *>(D + (d < nd ? d : 0) * ldl + pc + 2 * k2);
    }
  };
  if (p0 < p1) fetch(p0);
  for (int64_t pc = p0; pc < p1; pc += VK) {
    if (pc != p0) __syncthreads();
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      const int e = tid + kThreads * i, r = e / (VK / 2), k2 = e % (VK / 2);
      const bool ok = J0 + r < M;
      sL[r][2 * k2] = ok ? rl[i].x : 0.0;
      sL[r][2 * k2 + 1] = ok ? rl[i].y : 0.0;
    }
#pragma unroll
    for (int i = 0; i < NDD; ++i) {
      const int e = tid + kThreads * i, d = e / (VK / 2), k2 = e % (VK / 2);
      sD[d][2 * k2] = d < nd ? rd[i].x : 0.0;
      sD[d][2 * k2 + 1] = d < nd ? rd[i].y : 0.0;
    }
    __syncthreads();
    if (pc + VK < p1) fetch(pc + VK);
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int k = (ks * WPT + wk) * 4 + (lane >> 4);
      acc = mfma64(sL[16 * rt + (lane & 15)][k], sD[lane & 15][k], acc);
    }
  }
  if constexpr (WPT > 1) {
    __syncthreads();
    double *red = &sL[0][0];  // WPT-1 partial tiles per row tile: 4 x 64 doubles each
    if (wk > 0) {
#pragma unroll
      for (int r = 0; r < 4; ++r) red[((rt * (WPT - 1) + wk - 1) * 4 + r) * 64 + lane] = acc[r];
    }
    __syncthreads();
    if (wk > 0) return;
#pragma unroll
    for (int q = 0; q < WPT - 1; ++q)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] += red[((rt * (WPT - 1) + q) * 4 + r) * 64 + lane];
  }
  const int d = lane & 15;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int64_t J = J0 + 16 * rt + 4 * r + (lane >> 4);
    if (J < M && d < nd) vpart[(int64_t(blockIdx.y) * M + J) * nd + d] = acc[r];
  }
}

// Pass 2 as J^T = V^T L, L streamed from HBM straight into registers.
// grid = (ldl/512, nsplit).  Wave w of workgroup (x, s) owns the 128 pairs
// from 512x + 128w; per step of 4 vectors J..J+3, lane l (n = l%16, g = l/16)
// loads L[J + g][base + 32i + 2n .. + 1] for i < 4 (16 lanes cover 256
// contiguous bytes of a row), and MFMA (i, c) produces output columns
// base + 32i + 2n + c.  A operand: V[J + l/16][l%16].  Groups of U steps are
// fetched one group ahead.  jpart[(s*nd + d)*ldl + p].
template <int U>
__global__ __launch_bounds__(kThreads) void k_coulomb_j_stream(
    const double *__restrict__ L, int64_t ldl, int64_t M, const double *__restrict__ V, int nd,
    int64_t jchunk, double *__restrict__ jpart) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int n16 = lane & 15, g = lane >> 4;
  const int64_t pw = int64_t(blockIdx.x) * 512 + 128 * w + 2 * n16;
  const int64_t ja = int64_t(blockIdx.y) * jchunk;
  const int64_t jz = min(ja + jchunk, M);
  const int ngrp = int(jz > ja ? (jz - ja + 4 * U - 1) / (4 * U) : 0);
  v4d acc[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) acc[t] = v4d{0.0, 0.0, 0.0, 0.0};
  v2d la[U][4], lb[U][4];
  double va[U], vb[U];
  auto fetch = [&](int q, v2d(&x)[U][4], double(&v)[U]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t J = ja + int64_t(q) * 4 * U + 4 * u + g;
      const bool ok = J < jz;
      const double *__restrict__ row = L + (ok ? J : ja) * ldl + pw;
#pragma unroll
      for (int i = 0; i < 4; ++i)
        x[u][i] = ok ? __builtin_nontemporal_load(reinterpret_cast<const v2d *>(row + 32 * i))
                     : v2d{0.0, 0.0};
      v[u] = (ok && n16 < nd) ? V[J * nd + n16] : 0.0;
    }
  };
  auto body = [&](int q, v2d(&cur)[U][4], double(&vc)[U], v2d(&nxt)[U][4], double(&vn)[U]) {
    if (q + 1 < ngrp) fetch(q + 1, nxt, vn);
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        acc[2 * i] = mfma64(vc[u], cur[u][i].x, acc[2 * i]);
        acc[2 * i + 1] = mfma64(vc[u], cur[u][i].y, acc[2 * i + 1]);
      }
  };
  if (ngrp > 0) {
    fetch(0, la, va);
    for (int q = 0; q < ngrp; q += 2) {
      body(q, la, va, lb, vb);
      if (q + 1 < ngrp) body(q + 1, lb, vb, la, va);
    }
  }
  // acc[t][rr]: density 4rr + g, pair pw + t
#pragma unroll
  for (int rr = 0; rr < 4; ++rr) {
    const int m = 4 * rr + g;
    if (m < nd) {
      double *__restrict__ out = jpart + (int64_t(blockIdx.y) * nd + m) * ldl + pw;
#pragma unroll
      for (int i = 0; i < 4; ++i)
        *reinterpret_cast<v2d *>(out + 32 * i) = v2d{acc[2 * i][rr], acc[2 * i + 1][rr]};
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

namespace {

// Work split and scratch of one Coulomb build (both passes).
struct CoulombPlan {
  int nd = 0, ndt = 0;
  bool tile = false;                   // nd >= 3: the MFMA passes
  int nsplit1 = 0, nsplit2 = 0;
  int64_t nblk = 0, pchunk = 0, jchunk = 0;
  size_t bytes_d = 0, bytes_vpart = 0, bytes_v = 0, bytes_jpart = 0;
};

constexpr int kTileVM = 64, kTileVK = 64, kTileU = 4;  // MFMA passes (nd >= 3)

inline size_t piece(size_t bytes) { return size_t(round_up(int64_t(bytes), 256)); }

// Pass 2 split of the streaming J pass for ndt densities: about 1024
// workgroups of two segments, 2048 of one.
void plan_j_stream(const Context *c, int nd, int ndt, CoulombPlan *p) {
  const int64_t M = c->nchol, nseg = c->ldl / kSeg;
  const int64_t nxj = ceil_div(nseg, j_width(ndt));
  int nsplit2 = int(std::min<int64_t>(std::max<int64_t>(ceil_div(2048 / j_width(ndt), nxj), 1), M));
  p->jchunk = ceil_div(M, nsplit2);
  p->nsplit2 = int(ceil_div(M, p->jchunk));
  p->bytes_jpart = sizeof(double) * size_t(p->nsplit2) * nd * c->ldl;
}

CoulombPlan plan_coulomb(const Context *c, int nd) {
  CoulombPlan p;
  const int64_t M = c->nchol, ldl = c->ldl;
  p.nd = nd;
  p.bytes_d = sizeof(double) * size_t(nd) * ldl;
  p.bytes_v = sizeof(double) * size_t(M) * nd;
  if (nd >= 3) {
    p.tile = true;
    p.nblk = ceil_div(M, kTileVM);
    const int64_t nk = ldl / kTileVK;
    int nsplit1 = int(std::min<int64_t>(std::max<int64_t>(ceil_div(2048, p.nblk), 1), nk));
    p.pchunk = ceil_div(nk, nsplit1) * kTileVK;  // whole chunks: ldl % VK == 0
    p.nsplit1 = int(ceil_div(ldl, p.pchunk));
    const int64_t npc = ldl / 512;
    int nsplit2 = int(std::min<int64_t>(std::max<int64_t>(ceil_div(2048, npc), 1),
                                        ceil_div(M, 4 * kTileU)));
    p.jchunk = round_up(ceil_div(M, nsplit2), 4 * kTileU);
    p.nsplit2 = int(ceil_div(M, p.jchunk));
    p.bytes_jpart = sizeof(double) * size_t(p.nsplit2) * nd * ldl;
  } else {
    p.ndt = nd <= 1 ? 1 : 2;
    const int jb = 8;
    p.nblk = ceil_div(M, jb);
    const int64_t nseg = ldl / kSeg;
    int nsplit1 = int(std::min<int64_t>(std::max<int64_t>(ceil_div(2048, p.nblk), 1), nseg));
    p.pchunk = ceil_div(nseg, nsplit1) * kSeg;
    p.nsplit1 = int(ceil_div(ldl, p.pchunk));
    plan_j_stream(c, nd, p.ndt, &p);
  }
  p.bytes_vpart = sizeof(double) * size_t(p.nsplit1) * M * nd;
  return p;
}

// Where the passes put their scratch: the context's grow-only workspaces, or
// (arena set) consecutive 256-byte pieces of a fixed device buffer that is
// neither allocated nor freed here.
struct Scratch {
  double *arena = nullptr;
  size_t cap = 0, used = 0;  // bytes
};

int take(Context *c, Scratch &s, DevBuf &ws, size_t bytes, double **out) {
  if (!s.arena) {
    const int rc = ensure(c, ws, bytes);
    if (rc) return rc;
    *out = static_cast<double *>(ws.ptr);
    return MJK_OK;
  }
  if (s.used + piece(bytes) > s.cap)
    return set_error(c, MJK_ERR_ARG, "coulomb: scratch exceeds the fixed buffer");
  *out = s.arena + s.used / sizeof(double);
  s.used += piece(bytes);
  return MJK_OK;
}

// Pass 2 for one or two densities (ndt = 1, 2) or their padded counts:
// J = L V, then the fixed-order reduction of the vector slices into jlt.
int j_pass(Context *c, Scratch &s, const CoulombPlan &p, int nd, const double *V, double *jlt,
           int64_t ldj) {
  int rc;
  double *jpart;
  if ((rc = take(c, s, c->w->jpart, p.bytes_jpart, &jpart))) return rc;
  phase_begin(c);
  switch (p.ndt) {
    case 1: launch_j<1>(c, V, nd, jpart, p.nsplit2, p.jchunk); break;
    case 2: launch_j<2>(c, V, nd, jpart, p.nsplit2, p.jchunk); break;
    case 4: launch_j<4>(c, V, nd, jpart, p.nsplit2, p.jchunk); break;
    case 8: launch_j<8>(c, V, nd, jpart, p.nsplit2, p.jchunk); break;
    default: launch_j<16>(c, V, nd, jpart, p.nsplit2, p.jchunk); break;
  }
  phase_end(c, MJK_PHASE_COULOMB_J);
  phase_begin(c);
  hipLaunchKernelGGL(k_reduce_j, dim3(unsigned(ceil_div(c->np, 256)), unsigned(nd)), dim3(256), 0,
                     c->stream, jpart, c->ldl, c->np, nd, p.nsplit2, jlt, ldj);
  phase_end(c, MJK_PHASE_REDUCE);
  return hip_check(c, hipGetLastError(), "coulomb launch");
}

int coulomb_run(Context *c, Scratch &s, int32_t nd, const double *dlt, int64_t ldd, double *jlt,
                int64_t ldj, double *v_out) {
  if (nd < 1 || nd > 16) return set_error(c, MJK_ERR_ARG, "coulomb: nd must be in [1,16]");
  if (ldd < c->np || ldj < c->np) return set_error(c, MJK_ERR_ARG, "coulomb: ld < nbas(nbas+1)/2");
  if (c->nloaded != c->nchol)
    return set_error(c, MJK_ERR_SHORT_READ, "coulomb: not all Cholesky vectors were uploaded");
  const int64_t M = c->nchol, ldl = c->ldl;
  const CoulombPlan p = plan_coulomb(c, nd);
  int rc;
  double *Dp, *vpart, *V;
  // staging copy of the densities (zero padded to ldl)
  if ((rc = take(c, s, c->w->d, p.bytes_d, &Dp)) || (rc = take(c, s, c->w->vpart, p.bytes_vpart, &vpart)))
    return rc;
  if (v_out) V = v_out;
  else if ((rc = take(c, s, c->w->v, p.bytes_v, &V))) return rc;
  phase_begin(c);
  hipLaunchKernelGGL(k_pack_dens, dim3(unsigned(ceil_div(ldl, 256)), unsigned(nd)), dim3(256), 0,
                     c->stream, dlt, ldd, c->np, ldl, Dp);
  phase_end(c, MJK_PHASE_PACK);

  if (p.tile) {
    // ---- pass 1: V = L D^T, split over pairs to fill the device
    phase_begin(c);
    hipLaunchKernelGGL((k_coulomb_v_tile<kTileVM, kTileVK>), dim3(unsigned(p.nblk), unsigned(p.nsplit1)),
                       dim3(kThreads), 0, c->stream, c->L, ldl, M, Dp, nd, p.pchunk, vpart);
    phase_end(c, MJK_PHASE_COULOMB_V);
    phase_begin(c);
    hipLaunchKernelGGL(k_reduce_v, dim3(unsigned(ceil_div(M * nd, 256))), dim3(256), 0, c->stream,
                       vpart, M, nd, p.nsplit1, V);
    phase_end(c, MJK_PHASE_REDUCE);
    // ---- pass 2: J = L^T V, split over vectors
    double *jpart;
    if ((rc = take(c, s, c->w->jpart, p.bytes_jpart, &jpart))) return rc;
    phase_begin(c);
    hipLaunchKernelGGL(k_coulomb_j_stream<kTileU>, dim3(unsigned(ldl / 512), unsigned(p.nsplit2)),
                       dim3(kThreads), 0, c->stream, c->L, ldl, M, V, nd, p.jchunk, jpart);
    phase_end(c, MJK_PHASE_COULOMB_J);
    phase_begin(c);
    hipLaunchKernelGGL(k_reduce_j, dim3(unsigned(ceil_div(c->np, 256)), unsigned(nd)), dim3(256), 0,
                       c->stream, jpart, ldl, c->np, nd, p.nsplit2, jlt, ldj);
    phase_end(c, MJK_PHASE_REDUCE);
    return hip_check(c, hipGetLastError(), "coulomb launch");
  }

  // ---- pass 1: V = L^T D
  phase_begin(c);
  switch (p.ndt) {
    case 1: launch_v<8, 1>(c, Dp, nd, vpart, p.nsplit1, p.pchunk); break;
    default: launch_v<8, 2>(c, Dp, nd, vpart, p.nsplit1, p.pchunk); break;
  }
  phase_end(c, MJK_PHASE_COULOMB_V);
  phase_begin(c);
  hipLaunchKernelGGL(k_reduce_v, dim3(unsigned(ceil_div(M * nd, 256))), dim3(256), 0, c->stream,
                     vpart, M, nd, p.nsplit1, V);
  phase_end(c, MJK_PHASE_REDUCE);
  return j_pass(c, s, p, nd, V, jlt, ldj);
}

// ---- the allocation-free Coulomb build ----------------------------------
// Its scratch is one fixed device array per GPU, allocated with the code
// object and held until the process ends.  Host code uses the address that
// hipGetSymbolAddress returns for the current device, never the array's name
// (that is a host-side stand-in, not a device address).
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr int kMaxDevices = 64;
__device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

}  // namespace

// Device address of this GPU's scratch array, looked up once per device (with
// that device current) and cached.
static int scratch_device_ptr(Context *c, double **out) {
  static double *ptr[kMaxDevices] = {};
  static std::mutex mu;
  if (c->device < 0 || c->device >= kMaxDevices)
    return set_error(c, MJK_ERR_ARG, "coulomb_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[c->device]) {
    int cur = -1;
    int rc;
    if ((rc = hip_check(c, hipGetDevice(&cur), "hipGetDevice"))) return rc;
    if (cur != c->device)
      return set_error(c, MJK_ERR_STATE, "coulomb_noalloc: the context's device is not current");
    void *p = nullptr;
    if ((rc = hip_check(c, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)")))
      return rc;
    ptr[c->device] = static_cast<double *>(p);
  }
  *out = ptr[c->device];
  return MJK_OK;
}

int coulomb(Context *c, int32_t nd, const double *dlt, int64_t ldd, double *jlt, int64_t ldj,
            double *v_out) {
  Scratch s;  // the context's workspaces
  return coulomb_run(c, s, nd, dlt, ldd, jlt, ldj, v_out);
}

int coulomb_noalloc(Context *c, int32_t nd, const double *dlt, int64_t ldd, double *jlt, int64_t ldj,
                    double *v_out) {
  if (nd < 1 || nd > 16) return set_error(c, MJK_ERR_ARG, "coulomb: nd must be in [1,16]");
  const CoulombPlan p = plan_coulomb(c, nd);
  const size_t need = piece(p.bytes_d) + piece(p.bytes_vpart) + piece(p.bytes_jpart) +
                      (v_out ? 0 : piece(p.bytes_v));
  if (need > kScratchBytes)
    return set_error(c, MJK_ERR_ARG,
                     "coulomb_noalloc: this build needs " + std::to_string(need >> 20) +
                         " MiB of scratch, more than the fixed " +
                         std::to_string(kScratchBytes >> 20) + " MiB");
  Scratch s;
  int rc;
  if ((rc = scratch_device_ptr(c, &s.arena))) return rc;
  s.cap = kScratchBytes;
  return coulomb_run(c, s, nd, dlt, ldd, jlt, ldj, v_out);
}

namespace {
int coulomb_exchange_run(Context *c, const double *dlt, int64_t ldd, double *jlt, int64_t ldj,
                         int32_t nk, const int32_t *nocc, const double *const *corb, double *klt,
                         int64_t ldk, double *v_out) {
  if (ldd < c->np || ldj < c->np) return set_error(c, MJK_ERR_ARG, "coulomb: ld < nbas(nbas+1)/2");
  if (c->nloaded != c->nchol)
    return set_error(c, MJK_ERR_SHORT_READ, "coulomb: not all Cholesky vectors were uploaded");
  if (nk < 1 || !nocc) return set_error(c, MJK_ERR_ARG, "exchange: nk < 1");
  int64_t tot = 0;
  for (int d = 0; d < nk; ++d) tot += std::max(nocc[d], 0);
  int rc;
  if (tot == 0) {  // no half-transform to carry the V pass
    if ((rc = coulomb(c, 1, dlt, ldd, jlt, ldj, v_out))) return rc;
    return exchange(c, nk, nocc, corb, klt, ldk, MJK_PREC_FP64);
  }
  const int64_t M = c->nchol;
  const int ntmax = c->n_pad / 64;  // AO tiles of the narrowest half-transform tile
  if ((rc = ensure(c, c->w->vpart, sizeof(double) * size_t(ntmax) * M)) ||
      (rc = ensure(c, c->w->v, sizeof(double) * M)))
    return rc;
  double *vpart = static_cast<double *>(c->w->vpart.ptr);
  double *V = v_out ? v_out : static_cast<double *>(c->w->v.ptr);
  // the fused pass must not outlive this call on any return path
  struct FuseReset {
    Context *c;
    ~FuseReset() { c->fv = FuseV{}; }
  } fuse_reset{c};
  // the half-transform reads the caller's density at packed indices < nbas(nbas+1)/2
  // only (masked elements read index 0), so no padded copy is needed
  c->fv.p = dlt;
  c->fv.vpart = vpart;
  c->fv.m = M;
  c->fv.ntile = 0;
  if ((rc = exchange(c, nk, nocc, corb, klt, ldk, MJK_PREC_FP64))) return rc;
  const int ntile = c->fv.ntile;
  c->fv = FuseV{};
  if (ntile < 1 || ntile > ntmax)
    return set_error(c, MJK_ERR_ARG, "coulomb_exchange: the half-transform did not run");
  phase_begin(c);
  hipLaunchKernelGGL(k_reduce_v, dim3(unsigned(ceil_div(M, 256))), dim3(256), 0, c->stream, vpart,
                     M, 1, ntile, V);
  phase_end(c, MJK_PHASE_REDUCE);
  CoulombPlan p;
  p.ndt = 1;
  plan_j_stream(c, 1, 1, &p);
  Scratch s;  // the context's workspaces
  return j_pass(c, s, p, 1, V, jlt, ldj);
}

}  // namespace

int coulomb_exchange(Context *c, const double *dlt, int64_t ldd, double *jlt, int64_t ldj,
                     int32_t nk, const int32_t *nocc, const double *const *corb, double *klt,
                     int64_t ldk, double *v_out) {
  return coulomb_exchange_run(c, dlt, ldd, jlt, ldj, nk, nocc, corb, klt, ldk, v_out);
}

}  // namespace mjk
