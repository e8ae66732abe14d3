// spk_probe.hip -- the gather-throughput ceiling: a stream of (index,
// value) pairs read with nontemporal loads, each value multiplied by
// x[index] with the indices uniform over a window of x.  An SpMV whose rows
// scatter over all columns performs exactly this gather per nonzero, so
// the rate at the SpMV's x window (all of x, or one column panel) bounds
// it where HBM streaming does not.
#include "spk_internal.hpp"

namespace spk {
namespace {

constexpr int kProbeBlock = 512;
constexpr int kProbePer = 8;  // pairs per thread per chunk
constexpr int kProbeGrid = 2048;

__global__ void probe_fill_kernel(int64_t m, uint64_t window, int32_t* __restrict__ idx,
                                  double* __restrict__ v) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= m) return;
  uint64_t z = 0x5851F42D4C957F2Dull + (uint64_t)(k + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  idx[k] = (int32_t)(z % window);
  v[k] = 1.0;
}

__global__ __launch_bounds__(kProbeBlock) void probe_gather_kernel(
    const int32_t* __restrict__ idx, const double* __restrict__ v,
    const double* __restrict__ x, int64_t nchunks, double* __restrict__ out) {
  const int64_t chunk = (int64_t)kProbeBlock * kProbePer;
  double acc = 0.0;
  for (int64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const int64_t base = c * chunk + threadIdx.x;
    int32_t id[kProbePer];
    double a[kProbePer];
#pragma unroll
    for (int u = 0; u < kProbePer; ++u) {
      id[u] = __builtin_nontemporal_load(idx + base + u * kProbeBlock);
      a[u] = __builtin_nontemporal_load(v + base + u * kProbeBlock);
    }
#pragma unroll
    for (int u = 0; u < kProbePer; ++u) acc += a[u] * x[id[u]];
  }
  out[blockIdx.x * (int64_t)kProbeBlock + threadIdx.x] = acc;
}

}  // namespace
}  // namespace spk

using namespace spk;

extern "C" spk_status spk_gather_rate(int64_t window, int64_t count, int reps,
                                      double* gathers_per_s,
                                      double* stream_bytes_per_s) {
  return guarded([&] {
    const int64_t chunk = (int64_t)kProbeBlock * kProbePer;
    SPK_REQUIRE(window >= 1 && window < INT32_MAX, SPK_INVALID_ARGUMENT,
                "window must be in [1, 2^31)");
    SPK_REQUIRE(count >= chunk && count % chunk == 0, SPK_INVALID_ARGUMENT,
                "count must be a positive multiple of 4096");
    SPK_REQUIRE(reps >= 1 && gathers_per_s && stream_bytes_per_s,
                SPK_INVALID_ARGUMENT, "bad argument");
    hipStream_t s = stream();
    DevBuf<int32_t> idx((size_t)count);
    DevBuf<double> v((size_t)count), x((size_t)window),
        out((size_t)kProbeGrid * kProbeBlock);
    SPK_HIP_TRY(hipMemsetAsync(x.p, 0, (size_t)window * 8, s));
    hipLaunchKernelGGL(probe_fill_kernel, dim3((unsigned)((count + 255) / 256)),
                       dim3(256), 0, s, count, (uint64_t)window, idx.p, v.p);
    SPK_HIP_TRY(hipGetLastError());
    hipEvent_t e0, e1;
    SPK_HIP_TRY(hipEventCreate(&e0));
    SPK_HIP_TRY(hipEventCreate(&e1));
    float best = 0.0f;
    for (int r = 0; r <= reps; ++r) {  // the first launch is a warm-up
      SPK_HIP_TRY(hipEventRecord(e0, s));
      hipLaunchKernelGGL(probe_gather_kernel, dim3(kProbeGrid), dim3(kProbeBlock),
                         0, s, (const int32_t*)idx.p, (const double*)v.p,
                         (const double*)x.p, count / chunk, out.p);
      SPK_HIP_TRY(hipGetLastError());
      SPK_HIP_TRY(hipEventRecord(e1, s));
      SPK_HIP_TRY(hipEventSynchronize(e1));
      float ms = 0.0f;
      SPK_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
      if (r > 0 && (best == 0.0f || ms < best)) best = ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *gathers_per_s = (double)count / (best * 1e-3);
    *stream_bytes_per_s = 12.0 * (double)count / (best * 1e-3);
  });
}
