// Measured peaks of the device the context runs on: register-only MFMA and
// VALU loops, and a device-to-device streaming copy.  Reported by bench.py
// beside the published MI355X figures.
#include "mjk_internal.h"

namespace mjk {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
constexpr int kAcc = 8;
constexpr int kIters = 2000;
constexpr int kBlocks = 2048;

// The stores only happen for an impossible sum; they keep the loops alive.
__global__ __launch_bounds__(256) void k_peak_f64(double *sink, double seed) {
  v4d acc[kAcc];
  for (int i = 0; i < kAcc; ++i) acc[i] = v4d{0, 0, 0, 0};
  const double a = seed + threadIdx.x * 1e-9, b = seed - threadIdx.x * 1e-9;
  for (int it = 0; it < kIters; ++it)
#pragma unroll
    for (int i = 0; i < kAcc; ++i) acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[i], 0, 0, 0);
  double s = 0;
  for (int i = 0; i < kAcc; ++i) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
  if (s == -1.2345e300) sink[blockIdx.x] = s;
}
__global__ __launch_bounds__(256) void k_peak_f32(float *sink, float seed) {
  v4f acc[kAcc];
  for (int i = 0; i < kAcc; ++i) acc[i] = v4f{0, 0, 0, 0};
  const float a = seed + threadIdx.x * 1e-9f, b = seed - threadIdx.x * 1e-9f;
  for (int it = 0; it < kIters; ++it)
#pragma unroll
    for (int i = 0; i < kAcc; ++i) acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[i], 0, 0, 0);
  float s = 0;
  for (int i = 0; i < kAcc; ++i) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
  if (s == -1.2345e30f) sink[blockIdx.x] = s;
}
__global__ __launch_bounds__(256) void k_peak_bf16(float *sink, float seed) {
  v4f acc[kAcc];
  for (int i = 0; i < kAcc; ++i) acc[i] = v4f{0, 0, 0, 0};
  bf16x8 a, b;
  for (int e = 0; e < 8; ++e) {
    a[e] = __bf16(seed + e + threadIdx.x);
    b[e] = __bf16(seed - e);
  }
  for (int it = 0; it < kIters; ++it)
#pragma unroll
    for (int i = 0; i < kAcc; ++i) acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i], 0, 0, 0);
  float s = 0;
  for (int i = 0; i < kAcc; ++i) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
  if (s == -1.2345e30f) sink[blockIdx.x] = s;
}
__global__ __launch_bounds__(256) void k_peak_fma64(double *sink, double seed) {
  double x[8];
  for (int i = 0; i < 8; ++i) x[i] = seed + threadIdx.x + i;
  const double m = 0.999999, c = 1e-7;
  for (int it = 0; it < kIters; ++it)
#pragma unroll
    for (int i = 0; i < 8; ++i) x[i] = fma(x[i], m, c);
  double s = 0;
  for (int i = 0; i < 8; ++i) s += x[i];
  if (s == -1.2345e300) sink[blockIdx.x] = s;
}
__global__ void k_copy(const v2d *__restrict__ src, v2d *__restrict__ dst, int64_t n) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

}  // namespace
}  // namespace mjk

using namespace mjk;

extern "C" int mjk_measure_peak(mjk_context *c, int32_t kind, double *value) {
  if (!c || !value) return MJK_ERR_ARG;
  const int64_t copy_bytes = int64_t(1) << 30;
  int rc;
  if ((rc = ensure(c, c->w->misc, kind == 4 ? size_t(2 * copy_bytes) : sizeof(double) * kBlocks)))
    return rc;
  double *buf = static_cast<double *>(c->w->misc.ptr);
  hipEvent_t e0, e1;
  if ((rc = hip_check(c, hipEventCreate(&e0), "event")) || (rc = hip_check(c, hipEventCreate(&e1), "event")))
    return rc;
  auto launch = [&]() {
    switch (kind) {
      case 0: hipLaunchKernelGGL(k_peak_f64, dim3(kBlocks), dim3(256), 0, c->stream, buf, 1.0); break;
      case 1: hipLaunchKernelGGL(k_peak_f32, dim3(kBlocks), dim3(256), 0, c->stream, (float *)buf, 1.0f); break;
      case 2: hipLaunchKernelGGL(k_peak_bf16, dim3(kBlocks), dim3(256), 0, c->stream, (float *)buf, 1.0f); break;
      case 3: hipLaunchKernelGGL(k_peak_fma64, dim3(kBlocks), dim3(256), 0, c->stream, buf, 1.0); break;
      default:
        hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, c->stream, (const v2d *)buf,
                           (v2d *)(buf + copy_bytes / 8), copy_bytes / 16);
    }
  };
  if (kind < 0 || kind > 4) {
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return set_error(c, MJK_ERR_ARG, "measure_peak: kind must be 0..4");
  }
  launch();  // warm-up
  hipEventRecord(e0, c->stream);
  launch();
  hipEventRecord(e1, c->stream);
  rc = hip_check(c, hipEventSynchronize(e1), "peak sync");
  float ms = 0.f;
  if (!rc) rc = hip_check(c, hipEventElapsedTime(&ms, e0, e1), "elapsed");
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc) return rc;
  const double waves = double(kBlocks) * 4 * kIters * kAcc;
  double work;
  switch (kind) {
    case 0: case 1: work = waves * 16 * 16 * 4 * 2; break;          // flop
    case 2: work = waves * 16 * 16 * 32 * 2; break;                 // flop
    case 3: work = double(kBlocks) * 256 * kIters * 8 * 2; break;   // flop
    default: work = 2.0 * double(copy_bytes); break;                // bytes read + written
  }
  *value = work / (double(ms) * 1e-3) / 1e12;  // TFLOP/s or TB/s
  return hip_check(c, hipGetLastError(), "peak launch");
}
