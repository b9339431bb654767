// Calibration: matrix-core throughput from register-only chains of one MFMA
// instruction (no memory traffic) and the HBM read rate of a plain streaming
// kernel, the measured ceilings beside the published peaks the rooflines
// use.
#include <hip/hip_runtime.h>

#include <cstring>

#include "engine_internal.hpp"

using namespace gain;

namespace {

typedef int i4v __attribute__((ext_vector_type(4)));
typedef int i16v __attribute__((ext_vector_type(16)));
typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kChains = 8;   // independent accumulators per wave
constexpr int kUnroll = 32;  // rounds per loop trip (amortizes the copies
                             // the compiler makes around the loop)

// Every wave runs `trips` * kUnroll rounds of kChains independent products;
// the accumulators are stored only when `store` is set (never, in timing
// runs: it keeps the products live for the compiler).
__global__ __launch_bounds__(256) void k_peak_i8(int trips, int store,
                                                 int* out) {
  const int t = (int)threadIdx.x;
  const i4v a = {t, 1, 2, 3}, b = {3, t, 1, 2};
  i16v acc[kChains];
#pragma unroll
  for (int j = 0; j < kChains; ++j) acc[j] = i16v{};
  for (int it = 0; it < trips; ++it)
#pragma unroll
    for (int u = 0; u < kUnroll; ++u)
#pragma unroll
      for (int j = 0; j < kChains; ++j)
        acc[j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, acc[j], 0, 0, 0);
  if (store)
#pragma unroll
    for (int j = 0; j < kChains; ++j)
      out[blockIdx.x * 256 + t] += acc[j][0] + acc[j][15];
}

__global__ __launch_bounds__(256) void k_peak_f64(int trips, int store,
                                                  double* out) {
  const int t = (int)threadIdx.x;
  const double a = 1e-3 * t, b = 1.0 + 1e-4 * t;
  d4 acc[kChains];
#pragma unroll
  for (int j = 0; j < kChains; ++j) acc[j] = d4{0.0, 0.0, 0.0, 0.0};
  for (int it = 0; it < trips; ++it)
#pragma unroll
    for (int u = 0; u < kUnroll; ++u)
#pragma unroll
      for (int j = 0; j < kChains; ++j)
        acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[j], 0, 0, 0);
  if (store)
#pragma unroll
    for (int j = 0; j < kChains; ++j)
      out[blockIdx.x * 256 + t] += acc[j][0] + acc[j][3];
}

// Read-only HBM stream: grid-stride 16-byte non-temporal loads, eight per
// trip, summed so the loads stay live; one double written per thread.
typedef double dbl2 __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(256) void k_read_stream(
    const dbl2* __restrict__ p, int64_t n, double* __restrict__ out) {
  const int64_t tid = blockIdx.x * 256ll + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * 256;
  dbl2 s = {0.0, 0.0};
  int64_t i = tid;
  for (; i + 7 * stride < n; i += 8 * stride) {
    dbl2 v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      v[u] = __builtin_nontemporal_load(p + i + u * stride);
#pragma unroll
    for (int u = 0; u < 8; ++u) s += v[u];
  }
  for (; i < n; i += stride) s += p[i];
  out[tid] = s.x + s.y;
}

}  // namespace

extern "C" int gain_hbm_read_peak(gain_ctx* c, double* bytes_per_s) {
  return guarded([&] {
    need(c);
    GAIN_CHECK_ARG(bytes_per_s, "null argument");
    DeviceGuard g(c->device);
    const int64_t n = 1ll << 29;  // 8 GiB of 16-byte elements
    const int blocks = 8192;
    void* buf = nullptr;
    GAIN_HIP(hipMalloc(&buf, (size_t)n * 16));
    double* out = c->w_tmp.get<double>((size_t)blocks * 256);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.f;
    hipError_t err = hipMemsetAsync(buf, 0, (size_t)n * 16, c->stream);
    if (err == hipSuccess) err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    auto launch = [&] {
      hipLaunchKernelGGL(k_read_stream, dim3(blocks), dim3(256), 0,
                         c->stream, (const dbl2*)buf, n, out);
    };
    if (err == hipSuccess) {
      launch();  // warm
      err = hipEventRecord(e0, c->stream);
    }
    const int reps = 4;
    if (err == hipSuccess) {
      for (int r = 0; r < reps; ++r) launch();
      err = hipEventRecord(e1, c->stream);
    }
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(buf);
    if (err != hipSuccess)
      fail(GAIN_ERR_HIP, "hbm read peak: %s", hipGetErrorString(err));
    *bytes_per_s = (double)n * 16.0 * reps / (ms * 1e-3);
  });
}

extern "C" int gain_mfma_peak(gain_ctx* c, const char* kind,
                              double* ops_per_s) {
  return guarded([&] {
    need(c);
    GAIN_CHECK_ARG(kind && ops_per_s, "null argument");
    const bool i8 = std::strcmp(kind, "i8") == 0;
    GAIN_CHECK_ARG(i8 || std::strcmp(kind, "f64") == 0,
                   "kind must be \"i8\" or \"f64\"");
    DeviceGuard g(c->device);
    // four waves per SIMD on every CU
    const int blocks = 1024;
    const int trips = i8 ? 500 : 250;
    const double macs = i8 ? 32.0 * 32 * 32 : 16.0 * 16 * 4;
    void* out = c->w_tmp.get<double>((size_t)blocks * 256);
    hipEvent_t e0, e1;
    GAIN_HIP(hipEventCreate(&e0));
    GAIN_HIP(hipEventCreate(&e1));
    auto launch = [&](int n) {
      if (i8)
        hipLaunchKernelGGL(k_peak_i8, dim3(blocks), dim3(256), 0, c->stream,
                           n, 0, (int*)out);
      else
        hipLaunchKernelGGL(k_peak_f64, dim3(blocks), dim3(256), 0, c->stream,
                           n, 0, (double*)out);
      check_launch("mfma peak");
    };
    launch(10);  // warm
    GAIN_HIP(hipEventRecord(e0, c->stream));
    launch(trips);
    GAIN_HIP(hipEventRecord(e1, c->stream));
    GAIN_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    GAIN_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const double waves = blocks * 4.0;
    *ops_per_s = 2.0 * macs * kChains * kUnroll * (double)trips * waves /
                 (ms * 1e-3);
  });
}
