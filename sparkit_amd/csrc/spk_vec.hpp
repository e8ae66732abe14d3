// spk_vec.hpp -- fused vector kernels (BLAS-1) used by the device solvers.
//
// launch_map:    elementwise update, grid-stride.
// launch_reduce: elementwise update fused with W sums (accumulated in
//                double), reduced deterministically: fixed grid, per-thread
//                sequential partials, wavefront shuffle tree, one partial per
//                block, then a one-block finisher kernel that adds the
//                partials in index order and calls fin(totals) in thread 0 --
//                where the solvers update their device-resident scalars
//                (alpha, beta, convergence flag) so the host never has to
//                read them inside the iteration loop.
// Both take an optional `done` flag: when *done != 0 the kernels return at
// once (a solve that converged mid-chunk turns the rest of the queued chunk
// into no-ops).
#pragma once

#include "spk_internal.hpp"

namespace spk {

constexpr int kVecBlock = 256;

inline int vec_grid(int64_t n) {
  int64_t g = (n + kVecBlock - 1) / kVecBlock;
  int64_t cap = (int64_t)ctx().num_cus * 8;
  if (cap < 64) cap = 64;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return (int)g;
}

template <typename F>
__global__ __launch_bounds__(kVecBlock) void map_kernel(int64_t n, F f,
                                                        const int* done) {
  if (done && *done) return;
  for (int64_t i = blockIdx.x * (int64_t)kVecBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kVecBlock)
    f(i);
}

template <int W, typename F>
__global__ __launch_bounds__(kVecBlock) void reduce_kernel(int64_t n, F f,
                                                           const int* done,
                                                           double* partials) {
  __shared__ double red[kVecBlock / 64];
  if (done && *done) return;
  double acc[W];
#pragma unroll
  for (int k = 0; k < W; ++k) acc[k] = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kVecBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kVecBlock)
    f(i, acc);
#pragma unroll
  for (int k = 0; k < W; ++k) {
    double t = block_reduce_sum<double, kVecBlock>(acc[k], red);
    if (threadIdx.x == 0) partials[(size_t)blockIdx.x * W + k] = t;
  }
}

// Adds count partials (W interleaved sums) in a fixed order and hands the
// totals to fin in thread 0.
template <int W, typename Fin>
__global__ __launch_bounds__(kFinishBlock) void finish_kernel(
    const double* partials, int count, Fin fin, const int* done) {
  __shared__ double red[kFinishBlock / 64];
  __shared__ double tot[W];
  if (done && *done) return;
#pragma unroll
  for (int k = 0; k < W; ++k) {
    // eight independent accumulators keep eight loads in flight per thread;
    // the order of the additions depends only on count
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int i = threadIdx.x;
    for (; i + 7 * kFinishBlock < count; i += 8 * kFinishBlock) {
#pragma unroll
      for (int u = 0; u < 8; ++u) s[u] += partials[(size_t)(i + u * kFinishBlock) * W + k];
    }
    for (; i < count; i += kFinishBlock) s[0] += partials[(size_t)i * W + k];
    double t = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
    t = block_reduce_sum<double, kFinishBlock>(t, red);
    if (threadIdx.x == 0) tot[k] = t;
  }
  if (threadIdx.x == 0) fin((const double*)tot);
}

template <int W, typename Fin>
void launch_finish(const double* partials, int count, const Fin& fin,
                   const int* done) {
  hipLaunchKernelGGL((finish_kernel<W, Fin>), dim3(1), dim3(kFinishBlock), 0,
                     stream(), partials, count, fin, done);
  SPK_HIP_TRY(hipGetLastError());
}

template <typename F>
void launch_map(int64_t n, F f, const int* done = nullptr) {
  if (n <= 0) return;
  hipLaunchKernelGGL((map_kernel<F>), dim3(vec_grid(n)), dim3(kVecBlock), 0,
                     stream(), n, f, done);
  SPK_HIP_TRY(hipGetLastError());
}

template <int W, typename F, typename Fin>
void launch_reduce(int64_t n, F f, Fin fin, const int* done = nullptr) {
  int g = vec_grid(n);
  double* parts = reduce_partials((size_t)g * W);
  hipLaunchKernelGGL((reduce_kernel<W, F>), dim3(g), dim3(kVecBlock), 0,
                     stream(), n, f, done, parts);
  SPK_HIP_TRY(hipGetLastError());
  launch_finish<W>(parts, g, fin, done);
}

// Small device-resident block of scalars a solver keeps between kernels.
// It lives in a workspace slot, so a solve allocates nothing once warm.
template <typename S>
struct DevState {
  S* d = nullptr;
  explicit DevState(int slot) { d = (S*)workspace().get(slot, sizeof(S)); }
  void upload(const S& h) {
    SPK_HIP_TRY(hipMemcpyAsync(d, &h, sizeof(S), hipMemcpyHostToDevice, stream()));
    SPK_HIP_TRY(hipStreamSynchronize(stream()));  // h may be a temporary
  }
  S download() {
    S h;
    SPK_HIP_TRY(hipMemcpyAsync(&h, d, sizeof(S), hipMemcpyDeviceToHost, stream()));
    SPK_HIP_TRY(hipStreamSynchronize(stream()));
    return h;
  }
};

}  // namespace spk
