// FP64 Fock assembly and energy, kept as separate steps after J and K.
//
// Reference: the LT update FLT += J - ExFac*K of src/fock_util/cho_fscf.F90:
// 276,301 and the energy of src/scf/eneclc.F90:58,77,104
// (E = h.D + 1/2 G.D with the folded density D).
#include "mjk_internal.h"

namespace mjk {
namespace {

// flt[p] += fj*jlt[p] + fk*klt[p]  (either source may be null)
__global__ void k_fock_update(double *__restrict__ flt, const double *__restrict__ jlt, double fj,
                              const double *__restrict__ klt, double fk, int64_t np) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= np) return;
  double v = flt[p];
  if (jlt) v += fj * jlt[p];
  if (klt) v += fk * klt[p];
  flt[p] = v;
}

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 512;

__device__ inline double block_reduce(double x, double *lds) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) x += __shfl_down(x, off, kWave);
  if (lane == 0) lds[wave] = x;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kRedThreads / kWave; ++w) s += lds[w];
  return s;
}

// part[blockIdx.x] = sum over this block's grid-stride range of
// h*d + 1/2 g*d (h or g may be null).
__global__ __launch_bounds__(kRedThreads) void k_energy_part(const double *__restrict__ h,
                                                             const double *__restrict__ d,
                                                             const double *__restrict__ g,
                                                             int64_t np, double *__restrict__ part) {
  __shared__ double lds[kRedThreads / kWave];
  double s = 0.0;
  for (int64_t p = int64_t(blockIdx.x) * kRedThreads + threadIdx.x; p < np;
       p += int64_t(gridDim.x) * kRedThreads) {
    double t = 0.0;
    if (h) t += h[p];
    if (g) t += 0.5 * g[p];
    s += t * d[p];
  }
  s = block_reduce(s, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(kRedThreads) void k_energy_final(const double *__restrict__ part,
                                                              int nparts, double *__restrict__ e) {
  __shared__ double lds[kRedThreads / kWave];
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += kRedThreads) s += part[i];
  s = block_reduce(s, lds);
  if (threadIdx.x == 0) e[0] = s;
}

}  // namespace

int fock_update(Context *c, int32_t nd, double *flt, int64_t ldf, const double *jlt, int64_t ldj,
                const int32_t *jsrc, const double *fj, const double *klt, int64_t ldk,
                const double *fk) {
  if (nd < 1) return set_error(c, MJK_ERR_ARG, "fock_update: nd < 1");
  if (ldf < c->np) return set_error(c, MJK_ERR_ARG, "fock_update: ldf < nbas(nbas+1)/2");
  phase_begin(c);
  for (int d = 0; d < nd; ++d) {
    const double *js = (jlt && jsrc && jsrc[d] >= 0) ? jlt + jsrc[d] * ldj : nullptr;
    const double *ks = klt ? klt + d * ldk : nullptr;
    hipLaunchKernelGGL(k_fock_update, dim3(unsigned(ceil_div(c->np, 256))), dim3(256), 0,
                       c->stream, flt + d * ldf, js, fj ? fj[d] : 0.0, ks, fk ? fk[d] : 0.0,
                       c->np);
  }
  phase_end(c, MJK_PHASE_ASSEMBLY);
  return hip_check(c, hipGetLastError(), "fock_update launch");
}

int energy(Context *c, const double *h, const double *dlt, const double *g, double *e_out_dev) {
  int rc;
  if (!dlt || !e_out_dev) return set_error(c, MJK_ERR_ARG, "energy: null pointer");
  if ((rc = ensure(c, c->w->red, sizeof(double) * kRedBlocks))) return rc;
  double *part = static_cast<double *>(c->w->red.ptr);
  phase_begin(c);
  hipLaunchKernelGGL(k_energy_part, dim3(kRedBlocks), dim3(kRedThreads), 0, c->stream, h, dlt, g,
                     c->np, part);
  hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(kRedThreads), 0, c->stream, part, kRedBlocks,
                     e_out_dev);
  phase_end(c, MJK_PHASE_ASSEMBLY);
  return hip_check(c, hipGetLastError(), "energy launch");
}

}  // namespace mjk
