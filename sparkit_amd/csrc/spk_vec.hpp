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

// ---- one-kernel scalar hand-off ("prologue" kernels)
//
// Instead of a one-block finisher kernel after a reduction, every block of
// the NEXT kernel sums the reduction's partials itself, in one fixed order
// (the same in every block, so every block derives bit-identical scalars),
// and derives the scalars it needs.  The kernel boundary orders the
// partials' stores; block 0 may write solver state, but only state that no
// block of the same kernel reads (state read by the same kernel is kept in
// step-parity double buffers by the solvers).  Worth it when the partials
// are few (each block reads all of them): the solvers use it below 4096.
constexpr int kProMaxPartials = 4096;

// Elements a prologue kernel loads before it sums the partials (up to
// kProPrefetch per thread), so the partials' latency overlaps the loads.
constexpr int kProPrefetch = 4;

// Per-thread sums of `count` W-interleaved partials, in a fixed order: the
// first kProMaxPartials are loaded all at once (independent loads, one
// round trip) and added pairwise in registers; any beyond that follow in a
// strided tail.
template <int W, int BLOCK>
__device__ __forceinline__ void thread_sum_partials(const double* __restrict__ parts,
                                                    int count, double* v) {
  constexpr int J = kProMaxPartials / BLOCK;
  double a[J][W];
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int i = threadIdx.x + j * BLOCK;
#pragma unroll
    for (int k = 0; k < W; ++k) a[j][k] = i < count ? parts[(size_t)i * W + k] : 0.0;
  }
  double tail[W];
#pragma unroll
  for (int k = 0; k < W; ++k) tail[k] = 0.0;
  for (int i = threadIdx.x + J * BLOCK; i < count; i += BLOCK) {
#pragma unroll
    for (int k = 0; k < W; ++k) tail[k] += parts[(size_t)i * W + k];
  }
#pragma unroll
  for (int h = 1; h < J; h *= 2) {
#pragma unroll
    for (int j = 0; j + h < J; j += 2 * h) {
#pragma unroll
      for (int k = 0; k < W; ++k) a[j][k] += a[j + h][k];
    }
  }
#pragma unroll
  for (int k = 0; k < W; ++k) v[k] = a[0][k] + tail[k];
}

// Four- and eight-wide element packs for prologue kernels' loads.
template <typename T>
struct Pack4 {
  T a, b, c, d;
};
template <typename T>
struct Pack8 {
  T a, b, c, d, e, f, g, h;
};

// Prologue: the totals of WI-wide partials (and of WJ-wide partials from a
// second reduction, appended) go to pro(totals, scalars, is_block0) in
// thread 0 (up to 8 scalars, shared); then f(i, ld(i), acc, scalars) over
// the elements with WO sums written as this kernel's partials.  ld reads
// nothing the prologue decides; the first kProPrefetch loads of each
// thread are issued before the partials are read.  gate() false: the whole
// launch is a no-op.
template <int WI, int WJ, int WO, typename Gate, typename Pro, typename Ld, typename F>
__global__ __launch_bounds__(kVecBlock) void reduce_pro_kernel(
    int64_t n, const double* __restrict__ parts_in, int count_in,
    const double* __restrict__ parts_in2, int count_in2, Gate gate, Pro pro, Ld ld,
    F f, double* __restrict__ parts_out) {
  constexpr int NW = kVecBlock / 64;
  constexpr int WT = WI + WJ;
  __shared__ double red[(WT > 1 ? WT : 1) * NW];
  __shared__ double sc[8];
  if (!gate()) return;
  using E = decltype(ld((int64_t)0));
  const int64_t stride = (int64_t)gridDim.x * kVecBlock;
  const int64_t i0 = blockIdx.x * (int64_t)kVecBlock + threadIdx.x;
  E e[kProPrefetch];
#pragma unroll
  for (int j = 0; j < kProPrefetch; ++j) {
    const int64_t i = i0 + j * stride;
    if (i < n) e[j] = ld(i);
  }
  double v[WT];
  thread_sum_partials<WI, kVecBlock>(parts_in, count_in, v);
  if constexpr (WJ > 0) thread_sum_partials<WJ, kVecBlock>(parts_in2, count_in2, v + WI);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < WT; ++k) {
    const double t = wave_reduce_sum(v[k]);
    if (lane == 0) red[k * NW + wid] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot[WT];
#pragma unroll
    for (int k = 0; k < WT; ++k) {
      double t = red[k * NW];
#pragma unroll
      for (int w = 1; w < NW; ++w) t += red[k * NW + w];
      tot[k] = t;
    }
    pro((const double*)tot, sc, blockIdx.x == 0);
  }
  __syncthreads();
  double acc[WO > 0 ? WO : 1];
#pragma unroll
  for (int k = 0; k < (WO > 0 ? WO : 1); ++k) acc[k] = 0.0;
#pragma unroll
  for (int j = 0; j < kProPrefetch; ++j) {
    const int64_t i = i0 + j * stride;
    if (i < n) f(i, e[j], acc, (const double*)sc);
  }
  for (int64_t i = i0 + kProPrefetch * stride; i < n; i += stride)
    f(i, ld(i), acc, (const double*)sc);
  if constexpr (WO > 0) {
#pragma unroll
    for (int k = 0; k < WO; ++k) {
      double t = block_reduce_sum<double, kVecBlock>(acc[k], red);
      if (threadIdx.x == 0) parts_out[(size_t)blockIdx.x * WO + k] = t;
    }
  }
}

// Blocks per CU of a prologue kernel: every block re-reads all partials
// of the previous kernel, so the grid (and with it the number of partials
// the next prologue reads) is kept smaller than vec_grid's.
#ifndef SPK_PRO_BLOCKS_PER_CU
#define SPK_PRO_BLOCKS_PER_CU 4
#endif
inline int pro_grid(int64_t n) {
  int64_t g = (n + kVecBlock - 1) / kVecBlock;
  int64_t cap = (int64_t)ctx().num_cus * SPK_PRO_BLOCKS_PER_CU;
  if (cap < 64) cap = 64;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return (int)g;
}

// Returns the grid (= the number of WO-wide partials written to parts_out).
template <int WI, int WJ, int WO, typename Gate, typename Pro, typename Ld, typename F>
int launch_reduce_pro(int64_t n, const double* parts_in, int count_in,
                      const double* parts_in2, int count_in2, Gate gate, Pro pro,
                      Ld ld, F f, double* parts_out) {
  const int g = pro_grid(n);
  hipLaunchKernelGGL((reduce_pro_kernel<WI, WJ, WO, Gate, Pro, Ld, F>), dim3(g),
                     dim3(kVecBlock), 0, stream(), n, parts_in, count_in,
                     parts_in2, count_in2, gate, pro, ld, f, parts_out);
  SPK_HIP_TRY(hipGetLastError());
  return g;
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
